// se2p_kspace.hip -- Fourier-space part of the spectral Ewald sum.
//
// Behaviour follows mex_stokes_slp_kspace.cpp:8-335 and
// mex_stokes_dlp_kspace.cpp:10-406:
//   1. spread every source with a truncated Gaussian over the (P+1)^2 grid
//      nodes from its anchor (fast Gaussian gridding, ewald_tools.cpp:86-114
//      for the anchor);
//   2. forward 2D DFT of each grid;
//   3. Stokeslet or stresslet filter, zero the k = 0 mode;
//   4. inverse 2D DFT, keep the real (SLP) or imaginary (DLP) part;
//   5. interpolate back to the targets with the same Gaussian and scale by
//      (4 xi^2/eta)^2 h^2 / pi / (4 pi).
//
// GPU mapping:
//   * Spreading is a deterministic gather.  Sources are sorted by the 8x8
//     block of grid nodes holding their anchor.  A 256-thread workgroup owns a
//     16x16 tile of grid nodes (one node per thread, accumulators in
//     registers) and visits the anchor blocks whose support can reach the
//     tile.  Each batch of 64 sources gets its separable Gaussian factors
//     (P+1 in x, P+1 in y) computed into LDS by four threads per source; then
//     every thread adds the batch to its node in sorted order.  No atomics, so
//     the grids are bitwise reproducible.
//   * The DLP needs three grids, not four: the stresslet filter only uses
//     n2 f1 + n1 f2 (mex_stokes_dlp_kspace.cpp:242-250).
//   * rocFFT real-to-complex / complex-to-real transforms replace MATLAB's
//     complex fft2/ifft2.  The reference filters the full complex spectrum
//     and keeps Re (SLP) or Im (DLP) of the inverse.  On the Nyquist lines
//     its wavenumber convention (index M/2 -> +pi M/L) makes the filtered
//     spectrum non-Hermitian, so here the filter is replaced by its
//     Hermitian part F_s(q) = (F(q) + F(-q))/2 (SLP) or the anti-Hermitian
//     part (F(q) - F(-q))/2 times -i (DLP), which yields exactly
//     Re/Im of the reference's inverse transform.  The inverse transform's
//     1/(Mx My) is folded into the filter.
//   * Interpolation: one thread per target, (P+1)^2 gathers per grid.
#include "se2p_device.h"
#include "se2p_fsort.h"
#include "se2p_internal.h"

#include <algorithm>

namespace se2p {

constexpr int kTile = 16;     // spreading tile edge (grid nodes)
constexpr int kBin = 8;       // anchor bin edge (grid nodes)
constexpr int kBatch = 64;    // sources per LDS batch
constexpr int kMaxRanges = 64;

struct SpreadSrc {
    const int2 *anchor;      // wrapped anchor (ax, ay) in [0,Mx) x [0,My)
    const double *fgg;       // [3][ns]: ex*e4y, e3x, e3y
    const double *q;         // [nq][ns]
    int64_t ns;
};

// ---- per-source gridding data -------------------------------------------
// Anchor bin of a point: kBin x kBin blocks of grid nodes, x-major.
struct BinKeyFn {
    const double *p;
    double Lx, Ly, h;
    int P, Mx, My, nby;
    __device__ int operator()(int i) const
    {
        int mx, my;
        double px, py;
        se2p_anchor(p[2 * i], p[2 * i + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
        const int ax = ((mx % Mx) + Mx) % Mx, ay = ((my % My) + My) % My;
        return (ax / kBin) * nby + ay / kBin;
    }
};

__global__ void k_bin_keys(BinKeyFn kf, int n, int *__restrict__ keys)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = kf(i);
}

// FGG factors of mex_stokes_slp_kspace.cpp:95-99 with c = -2 xi^2 / eta.
__device__ inline void fgg_factors(double px, double py, double c, double w, double h,
                                   double *bx, double *e3x, double *e3y)
{
    double ex = exp(c * (px * px + py * py + 2 * w * px));
    double e4y = exp(2 * c * w * py);
    *bx = ex * e4y;
    *e3x = exp(-2 * c * h * px);
    *e3y = exp(-2 * c * h * py);
}

// Gridding data of input point o at sorted slot i: wrapped anchor, FGG
// factors and the densities of nr right-hand sides (kind 0: q = f, 2 comps;
// kind 1: q = (n1 f1, n2 f1 + n1 f2, n2 f2)).
template <int KIND>
__device__ __forceinline__ void gather_grid_source(const double *__restrict__ p,
                                                   const double *__restrict__ f,
                                                   const double *__restrict__ nv, int o, int i,
                                                   int n, int nreal, int nr, int64_t dstride,
                                                   double Lx, double Ly, double h, int P, int Mx,
                                                   int My, double c, double w,
                                                   int2 *__restrict__ anchor,
                                                   double *__restrict__ fgg,
                                                   double *__restrict__ q)
{
    int mx, my;
    double px, py;
    se2p_anchor(p[2 * o], p[2 * o + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
    anchor[i] = make_int2(((mx % Mx) + Mx) % Mx, ((my % My) + My) % My);
    double bx, e3x, e3y;
    fgg_factors(px, py, c, w, h, &bx, &e3x, &e3y);
    fgg[i] = bx;
    fgg[n + i] = e3x;
    fgg[2 * n + i] = e3y;
    // right-hand sides nreal..nr-1 only pad the last group: zero densities
    if (KIND == 0) {
        for (int r = 0; r < nr; ++r) {
            const double *fr = f + (int64_t)r * dstride;
            q[(int64_t)(2 * r) * n + i] = r < nreal ? fr[2 * o] : 0.0;
            q[(int64_t)(2 * r + 1) * n + i] = r < nreal ? fr[2 * o + 1] : 0.0;
        }
    } else {
        double n1 = nv[2 * o], n2 = nv[2 * o + 1];
        for (int r = 0; r < nr; ++r) {
            const double *fr = f + (int64_t)r * dstride;
            double f1 = r < nreal ? fr[2 * o] : 0.0, f2 = r < nreal ? fr[2 * o + 1] : 0.0;
            q[(int64_t)(3 * r) * n + i] = n1 * f1;
            q[(int64_t)(3 * r + 1) * n + i] = n2 * f1 + n1 * f2;
            q[(int64_t)(3 * r + 2) * n + i] = n2 * f2;
        }
    }
}

template <int KIND>
__global__ void k_gather_grid_src(const double *__restrict__ p, const double *__restrict__ f,
                                  const double *__restrict__ nv, const int *__restrict__ perm,
                                  int n, int nreal, int nr, int64_t dstride, double Lx, double Ly,
                                  double h, int P, int Mx, int My, double c, double w,
                                  int2 *__restrict__ anchor, double *__restrict__ fgg,
                                  double *__restrict__ q)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gather_grid_source<KIND>(p, f, nv, perm[i], i, n, nreal, nr, dstride, Lx, Ly, h, P, Mx, My, c,
                             w, anchor, fgg, q);
}

// The same gather run by the final scatter of the fused sort.
template <int KIND>
struct GridSrcGather {
    static constexpr bool kBox = false;
    const double *p, *f, *nv;
    int n, nreal, nr;
    int64_t dstride;
    double Lx, Ly, h;
    int P, Mx, My;
    double c, w;
    int2 *anchor;
    double *fgg, *q;
    __device__ void operator()(int i, int o, int) const
    {
        gather_grid_source<KIND>(p, f, nv, o, i, n, nreal, nr, dstride, Lx, Ly, h, P, Mx, My, c, w,
                                 anchor, fgg, q);
    }
};

// e1[k] = exp(c h^2 (k - P/2)^2), k = 0..P  (mex_stokes_slp_kspace.cpp:64-67)
__device__ inline void fill_e1(double *e1, int P, double ch2)
{
    for (int k = threadIdx.x; k <= P; k += blockDim.x) {
        int j = k - P / 2;
        e1[k] = exp(ch2 * j * j);
    }
}

// ---- spreading -------------------------------------------------------------
struct SpreadArgs {
    SpreadSrc s;
    const int *boff;   // bin offsets [nbx*nby + 1]
    int nbx, nby;
    int Mx, My, P;
    double ch2;        // c h^2
    double *grids;     // [nq][Mx][My]
    int64_t gstride;   // Mx*My
    // work split of the strip kernel: tile t owns chunks
    // [chunk_off[t], chunk_off[t+1]) of up to `chunk` candidate sources each
    const int *chunk_off;
    int ntx, nty, chunk;
    double *partial;   // [chunk id][256 nodes][nq] for tiles with several chunks
    const int *block_tile;                 // workgroup -> tile
    const int *rng_start, *rng_len;        // [tile][kTR] candidate runs
    int64_t max_chunks;                    // partial-buffer stride per group
};

// Launch arguments of up to kPack boxes, passed by value: box j owns the
// workgroups [off[j], off[j+1]) of the launch.
constexpr int kPack = 16;

struct SpreadPack {
    SpreadArgs a[kPack];
    int off[kPack + 1];
    int n;
};
static_assert(sizeof(SpreadPack) <= 4096, "kernel argument size");

// The box owning workgroup b (off ascending, off[0] = 0); wave-uniform.
__device__ __forceinline__ int pack_slot(const int *off, int n, int b)
{
    int j = 0;
    while (j + 1 < n && b >= off[j + 1]) ++j;
    return j;
}

// Anchor bins whose anchors a satisfy (t - a) mod M in [0, P] for some node t
// of [t0, t0 + kTile), i.e. anchors in the cyclic range [t0 - P, t0 + kTile - 1]:
// at most two runs of consecutive bins, (b0[0], n[0]) and (b0[1], n[1]), in
// cyclic order, each bin once.  Returns the number of runs.
__host__ __device__ inline int bin_runs(int t0, int M, int P, int nb, int *b0, int *n)
{
    const int span = kTile + P < M ? kTile + P : M;
    const int s = ((t0 - P) % M + M) % M;
    const int e = s + span - 1;   // last anchor, possibly >= M (wraps)
    if (e < M) {
        b0[0] = s / kBin;
        n[0] = e / kBin - s / kBin + 1;
        return 1;
    }
    const int b1 = s / kBin;
    b0[0] = b1;
    n[0] = nb - b1;
    const int e2 = (e - M) / kBin;
    const int n2 = (e2 + 1 < b1 ? e2 + 1 : b1);
    if (n2 <= 0) return 1;
    b0[1] = 0;
    n[1] = n2;
    return 2;
}

// Bin list in cyclic order: position p of the runs above.
__host__ __device__ inline int bin_at(const int *b0, const int *n, int p)
{
    return p < n[0] ? b0[0] + p : b0[1] + (p - n[0]);
}

// Per-tile candidate ranges of the bin-sorted sources: for x-bin position xp
// and y-run yr, slot xp*2 + yr holds (start, length) of a run of sources.
constexpr int kTR = 2 * 20;   // slots per tile (x-bins <= 20 for P <= 126)

template <int NQ, bool WRAP>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a)
{
    const int grp = blockIdx.z;   // group of right-hand sides
    extern __shared__ double lds[];
    const int P1 = a.P + 1;
    double *WX = lds;                       // [kBatch][P1]
    double *WY = WX + kBatch * P1;          // [kBatch][P1]
    double *Q = WY + kBatch * P1;           // [kBatch][NQ]
    double *E1 = Q + kBatch * NQ;           // [P1]
    int *AX = (int *)(E1 + P1);             // [kBatch]
    int *AY = AX + kBatch;                  // [kBatch]
    __shared__ int rstart[kMaxRanges], rpref[kMaxRanges + 1];
    __shared__ int nrange;

    const int tid = threadIdx.x;
    const int X0 = blockIdx.x * kTile, Y0 = blockIdx.y * kTile;
    fill_e1(E1, a.P, a.ch2);
    if (tid == 0) {
        int xb0[2], xn[2], yb0[2], yn[2];
        const int nxr = bin_runs(X0, a.Mx, a.P, a.nbx, xb0, xn);
        const int nyr = bin_runs(Y0, a.My, a.P, a.nby, yb0, yn);
        const int nbxl = xn[0] + (nxr > 1 ? xn[1] : 0);
        int nr = 0, tot = 0;
        for (int i = 0; i < nbxl; ++i) {
            const int bx = bin_at(xb0, xn, i);
            for (int r = 0; r < nyr; ++r) {
                const int key0 = bx * a.nby + yb0[r];
                const int s0 = a.boff[key0], s1 = a.boff[key0 + yn[r]];
                if (s1 > s0 && nr < kMaxRanges) {
                    rstart[nr] = s0;
                    rpref[nr] = tot;
                    tot += s1 - s0;
                    ++nr;
                }
            }
        }
        rpref[nr] = tot;
        nrange = nr;
    }
    __syncthreads();
    const int total = rpref[nrange];
    const int X = X0 + tid / kTile, Y = Y0 + tid % kTile;
    double acc[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) acc[c] = 0.0;

    const int half = (P1 + 1) / 2;
    for (int base = 0; base < total; base += kBatch) {
        // fill: 4 threads per source
        {
            const int sl = tid >> 2, part = tid & 3;
            const int v = base + sl;
            if (v < total) {
                int r = 0;
                while (rpref[r + 1] <= v) ++r;
                const int k = rstart[r] + (v - rpref[r]);
                const double bx = a.s.fgg[k], e3x = a.s.fgg[a.s.ns + k],
                             e3y = a.s.fgg[2 * a.s.ns + k];
                const int i0 = (part & 1) ? half : 0;
                const int i1 = (part & 1) ? P1 : half;
                const double step = (part < 2) ? e3x : e3y;
                double v0 = (part < 2) ? bx : 1.0;
                // v0 *= step^i0 by repeated squaring
                {
                    double pw = step, acc0 = 1.0;
                    for (int e = i0; e > 0; e >>= 1) {
                        if (e & 1) acc0 *= pw;
                        pw *= pw;
                    }
                    v0 *= acc0;
                }
                double *W = (part < 2) ? (WX + sl * P1) : (WY + sl * P1);
                for (int i = i0; i < i1; ++i) {
                    W[i] = v0 * E1[i];
                    v0 *= step;
                }
                if (part == 0) {
                    int2 an = a.s.anchor[k];
                    AX[sl] = an.x;
                    AY[sl] = an.y;
#pragma unroll
                    for (int c = 0; c < NQ; ++c)
                        Q[sl * NQ + c] = a.s.q[((int64_t)grp * NQ + c) * a.s.ns + k];
                }
            }
        }
        __syncthreads();
        const int nb = min(kBatch, total - base);
        for (int s = 0; s < nb; ++s) {
            int i = X - AX[s];
            i += (i < 0) ? a.Mx : 0;
            int j = Y - AY[s];
            j += (j < 0) ? a.My : 0;
            if (!WRAP) {
                if (i <= a.P && j <= a.P) {
                    const double wgt = WX[s * P1 + i] * WY[s * P1 + j];
#pragma unroll
                    for (int c = 0; c < NQ; ++c) acc[c] += wgt * Q[s * NQ + c];
                }
            } else {
                for (int ii = i; ii <= a.P; ii += a.Mx)
                    for (int jj = j; jj <= a.P; jj += a.My) {
                        const double wgt = WX[s * P1 + ii] * WY[s * P1 + jj];
#pragma unroll
                        for (int c = 0; c < NQ; ++c) acc[c] += wgt * Q[s * NQ + c];
                    }
            }
        }
        __syncthreads();
    }
    if (X < a.Mx && Y < a.My) {
        const int64_t idx = (int64_t)X * a.My + Y;
#pragma unroll
        for (int c = 0; c < NQ; ++c) a.grids[((int64_t)grp * NQ + c) * a.gstride + idx] = acc[c];
    }
}


// LDS writes of a wavefront visible to all its lanes (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l's value of v as a wave-uniform value (l uniform).
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Strip variant (grids with Mx > P and My > P + kPad): a lane owns a column
// strip of 4 nodes (X, Ys..Ys+3) of the 16x16 tile, so the x factor and the
// window test are shared by 4 nodes, and the y factors are zero-padded by
// kPad on both sides so the y overlap needs no test.  The candidate list of
// the tile is cut into batches of kSB sources; wavefront w takes batches
// w, w+4, ... into its own LDS area, and the four partial tiles are added in
// a fixed order at the end (deterministic).
constexpr int kPad = 3;

// sources per wavefront batch.  The four batch areas set the workgroup's LDS
// (about 30 KB at P = 24 with three grids, 36 KB with twelve).  With the tile
// update on the matrix cores, batches of 16 also for five and more grids
// (fewer fills and chunks) took the twelve-grid spreading of 64 densities on
// a 1e5-point, 512^2 workload from 4.70 to 3.15 ms.
__host__ __device__ constexpr int strip_batch(int) { return 16; }

__host__ __device__ inline size_t strip_lds_doubles(int P, int NQ)
{
    const size_t sb = strip_batch(NQ);
    const size_t P1 = P + 1, PW = P1 + 2 * kPad;
    const size_t per_wave = sb * (P1 + PW + NQ) + sb;   // + sb doubles holding 2*sb ints
    const size_t batch = 4 * per_wave + P1;
    const size_t red = 64 * 4 * (size_t)NQ;
    return batch > red ? batch : red;
}

// WIDE (Mx, My >= P + 16): no node of a tile is reached through the periodic
// wrap from the far side, so a source's reach over the tile is one interval
// per axis and the test is one unsigned compare; the source loop has no
// branch (a lane outside the support adds an exact zero), which lets the
// LDS reads of the next sources run ahead.
// One workgroup of the strip spreading: workgroup index b of the box's chunk
// list, right-hand-side group grp.
typedef double se2p_v4d __attribute__((ext_vector_type(4)));

// MF (WIDE only): the source loop on the matrix cores.  For a group of four
// sources g = 0..3 of the batch the tile update is the product
//   H[y][x] += sum_g wy_g(y) (q_g wx_g(x)),   y, x = 0..15 (tile offsets),
// one v_mfma_f64_16x16x4f64 per grid: lane l supplies A[y = l % 16][g = l / 16]
// (a y weight) and B[g = l / 16][x = l % 16] (a density times an x weight),
// and receives H[y = l / 16 + 4 r][x = l % 16] in accumulator r.  A weight
// outside the source's support, and every weight of a source past the
// batch's end, is zero.
template <int NQ, bool WIDE, bool MF>
__device__ __forceinline__ void spread_strip_wg(const SpreadArgs &a, const int b, const int grp)
{
    static_assert(!MF || WIDE, "the matrix-core loop needs one support interval per axis");
    constexpr int kSB = strip_batch(NQ);
    extern __shared__ double lds[];
    const int P1 = a.P + 1, PW = P1 + 2 * kPad;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t per_wave = (size_t)kSB * (P1 + PW + NQ) + kSB;
    double *WX = lds + wave * per_wave;       // [kSB][P1]
    double *WY = WX + kSB * P1;               // [kSB][PW]
    double *Q = WY + kSB * PW;                // [kSB][NQ]
    int *AXY = (int *)(Q + kSB * NQ);         // [kSB][2]
    double *E1 = lds + 4 * per_wave;          // [P1]
    __shared__ int rstart[kMaxRanges], rpref[kMaxRanges + 1];
    __shared__ int nrange;

    // grp: group of right-hand sides (grids NQ*grp .. NQ*grp+NQ-1)
    const double *qg = a.s.q + (int64_t)grp * NQ * a.s.ns;
    // block -> (tile, chunk of the tile's candidate list)
    const int ntiles = a.ntx * a.nty;
    if (b >= a.chunk_off[ntiles]) return;   // uniform over the workgroup
    const int tile = a.block_tile[b];
    const int chunk = b - a.chunk_off[tile];
    const int nchunk = a.chunk_off[tile + 1] - a.chunk_off[tile];
    const int X0 = (tile / a.nty) * kTile, Y0 = (tile % a.nty) * kTile;
    fill_e1(E1, a.P, a.ch2);
    if (WIDE)   // zero padding of the y rows, written once
        for (size_t i = threadIdx.x; i < 4 * per_wave; i += 256) lds[i] = 0.0;
    if (threadIdx.x < 64) {
        // candidate runs of the tile (precomputed), prefix sums by wave 0
        const int l = threadIdx.x;
        const int st = l < kTR ? a.rng_start[tile * kTR + l] : 0;
        const int ln = l < kTR ? a.rng_len[tile * kTR + l] : 0;
        int x = ln;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d);
            if (l >= d) x += y;
        }
        if (l < kTR) {
            rstart[l] = st;
            rpref[l + 1] = x;
        }
        if (l == 0) {
            rpref[0] = 0;
            nrange = kTR;
        }
    }
    __syncthreads();
    const int c0 = chunk * a.chunk;
    const int c1 = min(rpref[nrange], c0 + a.chunk);
    const int X = X0 + (lane & 15);
    const int Ys = Y0 + 4 * (lane >> 4);
    double acc[4][NQ];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc[k][c] = 0.0;
    se2p_v4d cm[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) cm[c] = se2p_v4d{0.0, 0.0, 0.0, 0.0};

    // Wavefront w takes batches w, w+4, ... of the chunk.  Each wave fills and
    // reads only its own batch area, so the waves run independently, with
    // wavefront-scope fences around the fill instead of workgroup barriers.
    const int nbatch = (c1 - c0 + kSB - 1) / kSB;
    int rcur = 0;   // candidate run holding this wave's next batch start (runs ascend)
    for (int bt = wave; bt < nbatch; bt += 4) {
        const int base = c0 + bt * kSB;
        while (rpref[rcur + 1] <= base) ++rcur;
        if (WIDE) {
            // Fill with every lane: LPR lanes per weight row (rows 0..kSB-1:
            // x rows, then the y rows), each lane one segment of the part of
            // the row that the tile uses (support nodes d .. d + kTile - 1
            // inside [0, P], d the source's tile offset on that axis),
            // starting from e3^j0 by squaring; the first lane of an x row
            // also writes the source's tile offsets and densities.  Row
            // entries outside that part keep older values: the product loops
            // read them only under a false support test, which selects 0.
            // The y rows' zero padding was written once at the start.
            constexpr int LPR = 64 / (2 * kSB);
            const int row = lane / LPR, part = lane % LPR;
            const bool isx = row < kSB;
            const int sl = isx ? row : row - kSB;
            const int v = base + sl;
            if (v < c1) {
                int r = rcur;
                while (rpref[r + 1] <= v) ++r;
                const int k = rstart[r] + (v - rpref[r]);
                const int2 an = a.s.anchor[k];
                // tile-relative offset d = (T0 - anchor) mod M taken in (P - M, P]:
                // grid line T0 + l is node l + d of the support
                const int M = isx ? a.Mx : a.My;
                int d = (isx ? X0 - an.x : Y0 - an.y);
                d += d < 0 ? M : 0;
                d = d <= a.P ? d : d - M;
                constexpr int seg = (kTile + LPR - 1) / LPR;
                const int hi = min(a.P + 1, d + kTile);
                const int j0 = max(d, 0) + part * seg, j1 = min(hi, j0 + seg);
                const double e3 = a.s.fgg[(isx ? 1 : 2) * a.s.ns + k];
                double w = isx ? a.s.fgg[k] : 1.0;
                double pw = e3;
                for (int e = j0; e; e >>= 1) {
                    if (e & 1) w *= pw;
                    pw *= pw;
                }
                double *dst = isx ? WX + sl * P1 : WY + sl * PW + kPad;
                for (int j = j0; j < j1; ++j) {
                    dst[j] = w * E1[j];
                    w *= e3;
                }
                if (isx && part == 0) {
                    int e = Y0 - an.y;
                    e += e < 0 ? a.My : 0;
                    AXY[2 * sl] = d;
                    AXY[2 * sl + 1] = e <= a.P ? e : e - a.My;
#pragma unroll
                    for (int c = 0; c < NQ; ++c) Q[sl * NQ + c] = qg[(int64_t)c * a.s.ns + k];
                }
            }
        } else if (lane < 2 * kSB) {
            const int sl = lane & (kSB - 1);
            const int v = base + sl;
            if (v < c1) {
                int r = rcur;
                while (rpref[r + 1] <= v) ++r;
                const int k = rstart[r] + (v - rpref[r]);
                if (lane < kSB) {
                    double w0 = a.s.fgg[k];
                    const double e3x = a.s.fgg[a.s.ns + k];
                    for (int i = 0; i < P1; ++i) {
                        WX[sl * P1 + i] = w0 * E1[i];
                        w0 *= e3x;
                    }
                    const int2 an = a.s.anchor[k];
                    AXY[2 * sl] = an.x;
                    AXY[2 * sl + 1] = an.y;
#pragma unroll
                    for (int c = 0; c < NQ; ++c) Q[sl * NQ + c] = qg[(int64_t)c * a.s.ns + k];
                } else {
                    double w0 = 1.0;
                    const double e3y = a.s.fgg[2 * a.s.ns + k];
                    double *row = WY + sl * PW;
                    for (int j = 0; j < kPad; ++j) row[j] = 0.0;
                    for (int j = 0; j < P1; ++j) {
                        row[kPad + j] = w0 * E1[j];
                        w0 *= e3y;
                    }
                    for (int j = kPad + P1; j < PW; ++j) row[j] = 0.0;
                }
            }
        }
        wave_lds_sync();
        const int nb = min(kSB, c1 - base);
        if constexpr (MF) {
            const int t = lane & 15, g = lane >> 4;
            for (int s0 = 0; s0 < nb; s0 += 4) {   // uniform trip count: every lane runs the MFMAs
                const int sg = s0 + g;
                const bool valid = sg < nb;
                const int sc = valid ? sg : 0;   // an LDS index inside the batch area
                const int2 dd = *(const int2 *)(AXY + 2 * sc);
                const int iu = t + dd.x, ju = t + dd.y;
                const bool hx = valid && (unsigned)iu <= (unsigned)a.P;
                const bool hy = (unsigned)ju <= (unsigned)a.P;
                const double wx = WX[sc * P1 + (hx ? iu : 0)];
                const double wy = WY[sc * PW + kPad + (hy ? ju : 0)];
                const double av = hy ? wy : 0.0;
                const double bx = hx ? wx : 0.0;
#pragma unroll
                for (int c = 0; c < NQ; ++c)
                    cm[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Q[sc * NQ + c] * bx, cm[c], 0, 0, 0);
            }
        } else if constexpr (WIDE && NQ <= 3) {
            // the batch's densities and tile offsets are the same for every
            // lane: lane s holds source s's, and v_readlane hands them to the
            // source loop as wave-uniform values (an LDS read of them would
            // return the same 8 bytes to all 64 lanes)
            const int lx = lane & 15, ly4 = 4 * (lane >> 4);
            const unsigned Pu = (unsigned)a.P, Pj = (unsigned)(a.P + kPad);
            double qv[NQ];
            int dxv = 0, dyv = 0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) qv[c] = 0.0;
            if (lane < kSB) {
#pragma unroll
                for (int c = 0; c < NQ; ++c) qv[c] = Q[lane * NQ + c];
                dxv = AXY[2 * lane];
                dyv = AXY[2 * lane + 1];
            }
            // constant trip count and a uniform guard: a loop holding
            // v_readlane (convergent) is unrolled only when fully unrollable
#pragma unroll
            for (int s = 0; s < kSB; ++s) {
                if (s < nb) {
                    const unsigned iu = (unsigned)(lx + __builtin_amdgcn_readlane(dxv, s));
                    const unsigned ju =
                        (unsigned)(ly4 + __builtin_amdgcn_readlane(dyv, s) + kPad);
                    const bool hit = iu <= Pu && ju <= Pj;
                    const double wx = WX[s * P1 + min(iu, Pu)];
                    const double wxe = hit ? wx : 0.0;
                    const double *wy = WY + s * PW + min(ju, Pj);   // zero-padded row
                    double q[NQ];
#pragma unroll
                    for (int c = 0; c < NQ; ++c) q[c] = readlane_f64(qv[c], s);
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const double wgt = wxe * wy[kk];
#pragma unroll
                        for (int c = 0; c < NQ; ++c) acc[kk][c] += wgt * q[c];
                    }
                }
            }
        } else if (WIDE) {
            const int lx = lane & 15, ly4 = 4 * (lane >> 4);
            const unsigned Pu = (unsigned)a.P, Pj = (unsigned)(a.P + kPad);
#pragma unroll 4
            for (int s = 0; s < nb; ++s) {
                const unsigned iu = (unsigned)(lx + AXY[2 * s]);
                const unsigned ju = (unsigned)(ly4 + AXY[2 * s + 1] + kPad);
                const bool hit = iu <= Pu && ju <= Pj;
                const double wx = WX[s * P1 + min(iu, Pu)];
                const double wxe = hit ? wx : 0.0;
                const double *wy = WY + s * PW + min(ju, Pj);   // zero-padded row
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double wgt = wxe * wy[kk];
#pragma unroll
                    for (int c = 0; c < NQ; ++c) acc[kk][c] += wgt * Q[s * NQ + c];
                }
            }
        } else {
#pragma unroll 2
        for (int s = 0; s < nb; ++s) {
            int i = X - AXY[2 * s];
            i += (i < 0) ? a.Mx : 0;
            int j0 = Ys - AXY[2 * s + 1];
            j0 += (j0 < -kPad) ? a.My : 0;
            if (i <= a.P && j0 <= a.P) {
                const double wx = WX[s * P1 + i];
                const double *wy = WY + s * PW + kPad + j0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double wgt = wx * wy[k];
#pragma unroll
                    for (int c = 0; c < NQ; ++c) acc[k][c] += wgt * Q[s * NQ + c];
                }
            }
        }
        }
        wave_lds_sync();   // reads of this batch done before the next fill
    }
    if constexpr (MF) {
        // accumulator r of lane l holds node (x, y) = (l % 16, l / 16 + 4 r)
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < NQ; ++c) acc[k][c] = cm[c][k];
    }
    __syncthreads();   // every wave is done with its batch area (reused below)
    // fixed-order reduction of the four wavefronts' tiles, one wave at a time
    double *red = lds;   // [64][4*NQ]; the batch areas are dead now
    for (int w = 1; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int c = 0; c < NQ; ++c) red[lane * 4 * NQ + k * NQ + c] = acc[k][c];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int c = 0; c < NQ; ++c) acc[k][c] += red[lane * 4 * NQ + k * NQ + c];
        }
        __syncthreads();
    }
    if (wave != 0) return;
    double *gg = a.grids + (int64_t)grp * NQ * a.gstride;
    if (MF && nchunk == 1) {
        if (X < a.Mx) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int Y = Y0 + (lane >> 4) + 4 * k;
                if (Y >= a.My) break;
                const int64_t idx = (int64_t)X * a.My + Y;
#pragma unroll
                for (int c = 0; c < NQ; ++c) gg[c * a.gstride + idx] = acc[k][c];
            }
        }
    } else if (MF) {
        // partial tile in the layout k_spread_combine reads (node (x, y) at
        // lane x + 16 (y / 4), slot y % 4)
        double *pt = a.partial + ((int64_t)grp * a.max_chunks + b) * 256 * NQ;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < NQ; ++c) pt[(c * 4 + (lane >> 4)) * 64 + (lane & 15) + 16 * k] = acc[k][c];
    } else if (nchunk == 1) {
        if (X < a.Mx) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int Y = Ys + k;
                if (Y >= a.My) break;
                const int64_t idx = (int64_t)X * a.My + Y;
#pragma unroll
                for (int c = 0; c < NQ; ++c) gg[c * a.gstride + idx] = acc[k][c];
            }
        }
    } else {
        // partial tile, added to the others in chunk order by k_spread_combine
        double *pt = a.partial + ((int64_t)grp * a.max_chunks + b) * 256 * NQ;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < NQ; ++c) pt[(c * 4 + k) * 64 + lane] = acc[k][c];
    }
}

// Strip spreading on grids with Mx, My >= P + 16 uses the matrix cores.
constexpr bool kSpreadMfma = true;

// One-dimensional grid of max_chunks x groups, groups fastest: a chunk's
// groups of grids are dispatched together, so with several groups (batched
// densities) the launch does not end on the heavy tiles of the last group.
template <int NQ, bool WIDE>
__global__ __launch_bounds__(256) void k_spread_strip(SpreadArgs a)
{
    const int ngrp = (int)(gridDim.x / (unsigned)a.max_chunks);
    const int b = (int)blockIdx.x / ngrp;
    spread_strip_wg<NQ, WIDE, WIDE && kSpreadMfma>(a, b, (int)blockIdx.x - b * ngrp);
}

// Many boxes in one launch: box j owns workgroups [off[j], off[j+1]) and each
// runs exactly the workgroup a single-box launch would run.
template <int NQ, bool WIDE>
__global__ __launch_bounds__(256) void k_spread_strip_boxes(SpreadPack p)
{
    const int j = pack_slot(p.off, p.n, blockIdx.x);
    spread_strip_wg<NQ, WIDE, WIDE && kSpreadMfma>(p.a[j], blockIdx.x - p.off[j], blockIdx.y);
}

// Tiles whose candidates were split over several workgroups: add the partial
// tiles in chunk order (deterministic) and write the grid.
template <int NQ>
__device__ __forceinline__ void spread_combine_wg(const SpreadArgs &a, const int tile, const int grp)
{
    double *gg = a.grids + (int64_t)grp * NQ * a.gstride;
    const int b0 = a.chunk_off[tile], b1 = a.chunk_off[tile + 1];
    if (b1 - b0 <= 1) return;
    const int lane = threadIdx.x;
    const int X = (tile / a.nty) * kTile + (lane & 15);
    const int Ys = (tile % a.nty) * kTile + 4 * (lane >> 4);
    if (X >= a.Mx) return;
    double acc[4][NQ];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc[k][c] = 0.0;
    for (int b = b0; b < b1; ++b) {
        const double *pt = a.partial + ((int64_t)grp * a.max_chunks + b) * 256 * NQ;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < NQ; ++c) acc[k][c] += pt[(c * 4 + k) * 64 + lane];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int Y = Ys + k;
        if (Y >= a.My) break;
        const int64_t idx = (int64_t)X * a.My + Y;
#pragma unroll
        for (int c = 0; c < NQ; ++c) gg[c * a.gstride + idx] = acc[k][c];
    }
}

template <int NQ>
__global__ __launch_bounds__(64) void k_spread_combine(SpreadArgs a)
{
    spread_combine_wg<NQ>(a, blockIdx.x, blockIdx.y);
}

// Box j owns workgroups [off[j], off[j+1]), one per tile.
template <int NQ>
__global__ __launch_bounds__(64) void k_spread_combine_boxes(SpreadPack p)
{
    const int j = pack_slot(p.off, p.n, blockIdx.x);
    spread_combine_wg<NQ>(p.a[j], blockIdx.x - p.off[j], blockIdx.y);
}

// Work split per tile (one wavefront per tile): the candidate runs of the
// bin-sorted sources, the number of spreading chunks (at least one per tile,
// so every grid node is written) and of interpolation chunks of 256 targets.
__device__ __forceinline__ void tile_work(const SpreadArgs &a, const int *__restrict__ tboff,
                                          int tile, int l, int *__restrict__ scnt,
                                          int *__restrict__ icnt, int *__restrict__ rng_start,
                                          int *__restrict__ rng_len)
{
    const int tx = tile / a.nty, ty = tile % a.nty;
    int xb0[2], xn[2], yb0[2], yn[2];
    const int nxr = bin_runs(tx * kTile, a.Mx, a.P, a.nbx, xb0, xn);
    const int nyr = bin_runs(ty * kTile, a.My, a.P, a.nby, yb0, yn);
    const int nbxl = xn[0] + (nxr > 1 ? xn[1] : 0);
    int st = 0, ln = 0;
    if (l < kTR) {
        const int xp = l >> 1, yr = l & 1;
        if (xp < nbxl && yr < nyr) {
            const int key0 = bin_at(xb0, xn, xp) * a.nby + yb0[yr];
            st = a.boff[key0];
            ln = a.boff[key0 + yn[yr]] - st;
        }
        rng_start[tile * kTR + l] = st;
        rng_len[tile * kTR + l] = ln;
    }
    int tot = ln;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) tot += __shfl_xor(tot, d);
    if (l == 0) {
        scnt[tile] = max(1, (tot + a.chunk - 1) / a.chunk);
        // targets of the tile: bins (2tx..2tx+1, 2ty..2ty+1)
        const int bx0 = 2 * tx, by0 = 2 * ty;
        const int nyb = min(2, a.nby - by0);
        int nt = tboff[bx0 * a.nby + by0 + nyb] - tboff[bx0 * a.nby + by0];
        if (bx0 + 1 < a.nbx)
            nt += tboff[(bx0 + 1) * a.nby + by0 + nyb] - tboff[(bx0 + 1) * a.nby + by0];
        icnt[tile] = (nt + 255) / 256;
    }
}

__global__ __launch_bounds__(64) void k_tile_work(SpreadArgs a, const int *__restrict__ tboff,
                                                  int *__restrict__ scnt, int *__restrict__ icnt,
                                                  int *__restrict__ rng_start,
                                                  int *__restrict__ rng_len)
{
    tile_work(a, tboff, blockIdx.x, threadIdx.x, scnt, icnt, rng_start, rng_len);
}

// The tile work as extra workgroups of the fused sort: one tile per wavefront.
struct TileWorkExtra {
    SpreadArgs a;
    const int *tboff;
    int *scnt, *icnt, *rng_start, *rng_len;
    int ntiles;
    int blocks(int nw) const { return ntiles > 0 ? (ntiles + nw - 1) / nw : 0; }
    template <int NW>
    __device__ void run(int xb, int *) const
    {
        const int tile = xb * NW + (int)(threadIdx.x >> 6);
        if (tile < ntiles) tile_work(a, tboff, tile, threadIdx.x & 63, scnt, icnt, rng_start, rng_len);
    }
};

// ---- filter ----------------------------------------------------------------
struct FilterArgs {
    double2 *spec;     // [ngrid][Mx][My/2+1]
    int64_t sstride;   // Mx*(My/2+1)
    int Mx, My, nrhs;
    double Lx, Ly, xi, eta, scale;   // scale = 1/(Mx My)
};

__device__ inline double wavenumber(int j, int M, double L)
{
    return (j <= M / 2) ? 2.0 * SE2P_PI / L * j : 2.0 * SE2P_PI / L * (j - M);
}

// Stokeslet filter coefficients at half-spectrum index (j, k), the inverse
// transform's 1/(Mx My) included (mex_stokes_slp_kspace.cpp:200): e times
// the matrix (a11, a12; a12, a22); a12 is dropped where exactly one index is
// on its Nyquist line (the filter's Hermitian part).  Zero at (0, 0).
struct SlpCoef {
    double e, a11, a22, a12;
};
__device__ inline SlpCoef slp_coef(const FilterArgs &a, int j, int k)
{
    SlpCoef c{0.0, 0.0, 0.0, 0.0};
    if (j == 0 && k == 0) return c;
    const double k1 = wavenumber(j, a.Mx, a.Lx);
    const double k2 = 2.0 * SE2P_PI / a.Ly * k;
    const double K2 = k1 * k1 + k2 * k2;
    c.e = (1.0 / (K2 * K2) + 0.25 / (K2 * a.xi * a.xi)) *
          exp(-0.25 * (1 - a.eta) / (a.xi * a.xi) * K2) * a.scale;
    const bool nyx = (2 * j == a.Mx), nyy = (2 * k == a.My);
    c.a11 = k2 * k2;
    c.a22 = k1 * k1;
    c.a12 = (nyx != nyy) ? 0.0 : k1 * k2;
    return c;
}

__device__ __forceinline__ void slp_apply(const SlpCoef &c, double2 h1, double2 h2, double2 &o1,
                                          double2 &o2)
{
    o1 = make_double2((c.a11 * h1.x - c.a12 * h2.x) * c.e, (c.a11 * h1.y - c.a12 * h2.y) * c.e);
    o2 = make_double2((c.a22 * h2.x - c.a12 * h1.x) * c.e, (c.a22 * h2.y - c.a12 * h1.y) * c.e);
}

// Stresslet filter coefficients at half-spectrum index (j, k), the inverse
// transform's 1/(Mx My) included (mex_stokes_dlp_kspace.cpp:232-247).  Only the
// anti-Hermitian part is kept: terms odd in k1 vanish on the x Nyquist line,
// terms odd in k2 on the y Nyquist line.
struct DlpCoef {
    double et, b111, b112, b122, b222;
};
__device__ inline DlpCoef dlp_coef(const FilterArgs &a, int j, int k)
{
    DlpCoef c{0.0, 0.0, 0.0, 0.0, 0.0};
    if (j == 0 && k == 0) return c;
    const double k1 = wavenumber(j, a.Mx, a.Lx);
    const double k2 = 2.0 * SE2P_PI / a.Ly * k;
    const double K2 = k1 * k1 + k2 * k2;
    const double e = exp(-0.25 * (1 - a.eta) / (a.xi * a.xi) * K2) / K2 / K2;
    const double t = 1 + 0.25 * K2 / a.xi / a.xi;
    c.et = e * t * a.scale;
    if (2 * j != a.Mx) {
        c.b111 = 3.0 * k1 * K2 - 2.0 * k1 * k1 * k1;
        c.b122 = k1 * K2 - 2.0 * k1 * k2 * k2;
    }
    if (2 * k != a.My) {
        c.b112 = k2 * K2 - 2.0 * k1 * k1 * k2;
        c.b222 = 3.0 * k2 * K2 - 2.0 * k2 * k2 * k2;
    }
    return c;
}

// The stresslet filter on one spectrum value of A = n1f1, B = n2f1+n1f2,
// C = n2f2: two outputs multiplied by -i, so that the real inverse transform
// yields the imaginary part of the reference's inverse.
__device__ __forceinline__ void dlp_apply(const DlpCoef &c, double2 A, double2 B, double2 C,
                                          double2 &o1, double2 &o2)
{
    const double o1r = c.et * (c.b111 * A.x + c.b112 * B.x + c.b122 * C.x);
    const double o1i = c.et * (c.b111 * A.y + c.b112 * B.y + c.b122 * C.y);
    const double o2r = c.et * (c.b112 * A.x + c.b122 * B.x + c.b222 * C.x);
    const double o2i = c.et * (c.b112 * A.y + c.b122 * B.y + c.b222 * C.y);
    o1 = make_double2(o1i, -o1r);
    o2 = make_double2(o2i, -o2r);
}

// Its transpose: the filter is a real convolution with an odd kernel, so the
// transpose multiplies by the complex conjugate of -i et b, i.e. +i et b,
// with the 2x3 coefficient matrix transposed (two inputs, three outputs).
__device__ __forceinline__ void dlp_adj_apply(const DlpCoef &c, double2 Z1, double2 Z2,
                                              double2 &o1, double2 &o2, double2 &o3)
{
    const double o1r = c.et * (c.b111 * Z1.x + c.b112 * Z2.x);
    const double o1i = c.et * (c.b111 * Z1.y + c.b112 * Z2.y);
    const double o2r = c.et * (c.b112 * Z1.x + c.b122 * Z2.x);
    const double o2i = c.et * (c.b112 * Z1.y + c.b122 * Z2.y);
    const double o3r = c.et * (c.b122 * Z1.x + c.b222 * Z2.x);
    const double o3i = c.et * (c.b122 * Z1.y + c.b222 * Z2.y);
    o1 = make_double2(-o1i, o1r);
    o2 = make_double2(-o2i, o2r);
    o3 = make_double2(-o3i, o3r);
}

// The filters as maps from input spectra to output spectra at one index
// (j, k) of the half spectrum, per right-hand side r: input grid in(r, c),
// c < nin, output grid out(r, o), o < nout, all in the same spectrum array.
//   FILTER_SLP      [2r], [2r+1]         -> [2r], [2r+1]
//   FILTER_DLP      [3r], [3r+1], [3r+2] -> [2r], [2r+1]  (A = n1f1,
//                   B = n2f1 + n1f2, C = n2f2)
//   FILTER_DLP_ADJ  [0], [1] (spectra of the spread du) -> [0], [1], [2]
//                   (the three stresslet densities), one right-hand side
//   FILTER_BOTH_*   forward and transposed filters of one spreading of
//                   2 + 2 (Stokeslet) or 3 + 2 (stresslet) grids at the same
//                   points: the densities' spectra [0, ncomp) filtered into
//                   [0, 2), the spread du's [ncomp, ncomp + 2) by the
//                   transposed filter (the Stokeslet filter is its own
//                   transpose) into [2, 2 + ncomp); one right-hand side.
// Every output of r is a grid no later right-hand side reads (2r + 1 <
// 3r' for r' > r), so right-hand sides may run one after the other in place.
enum FilterKind { FILTER_SLP, FILTER_DLP, FILTER_DLP_ADJ, FILTER_BOTH_SLP, FILTER_BOTH_DLP };

template <int FK>
struct FilterOp;
template <>
struct FilterOp<FILTER_SLP> {
    static constexpr int nin = 2, nout = 2;
    __device__ static int in(int r, int c) { return 2 * r + c; }
    __device__ static int out(int r, int o) { return 2 * r + o; }
    __device__ static SlpCoef coef(const FilterArgs &a, int j, int k) { return slp_coef(a, j, k); }
    __device__ static void apply(const SlpCoef &c, const double2 *x, double2 *y)
    {
        slp_apply(c, x[0], x[1], y[0], y[1]);
    }
};
template <>
struct FilterOp<FILTER_DLP> {
    static constexpr int nin = 3, nout = 2;
    __device__ static int in(int r, int c) { return 3 * r + c; }
    __device__ static int out(int r, int o) { return 2 * r + o; }
    __device__ static DlpCoef coef(const FilterArgs &a, int j, int k) { return dlp_coef(a, j, k); }
    __device__ static void apply(const DlpCoef &c, const double2 *x, double2 *y)
    {
        dlp_apply(c, x[0], x[1], x[2], y[0], y[1]);
    }
};
template <>
struct FilterOp<FILTER_DLP_ADJ> {
    static constexpr int nin = 2, nout = 3;
    __device__ static int in(int, int c) { return c; }
    __device__ static int out(int, int o) { return o; }
    __device__ static DlpCoef coef(const FilterArgs &a, int j, int k) { return dlp_coef(a, j, k); }
    __device__ static void apply(const DlpCoef &c, const double2 *x, double2 *y)
    {
        dlp_adj_apply(c, x[0], x[1], y[0], y[1], y[2]);
    }
};
template <>
struct FilterOp<FILTER_BOTH_SLP> {
    static constexpr int nin = 4, nout = 4;
    __device__ static int in(int, int c) { return c; }
    __device__ static int out(int, int o) { return o; }
    __device__ static SlpCoef coef(const FilterArgs &a, int j, int k) { return slp_coef(a, j, k); }
    __device__ static void apply(const SlpCoef &c, const double2 *x, double2 *y)
    {
        slp_apply(c, x[0], x[1], y[0], y[1]);
        slp_apply(c, x[2], x[3], y[2], y[3]);
    }
};
template <>
struct FilterOp<FILTER_BOTH_DLP> {
    static constexpr int nin = 5, nout = 5;
    __device__ static int in(int, int c) { return c; }
    __device__ static int out(int, int o) { return o; }
    __device__ static DlpCoef coef(const FilterArgs &a, int j, int k) { return dlp_coef(a, j, k); }
    __device__ static void apply(const DlpCoef &c, const double2 *x, double2 *y)
    {
        dlp_apply(c, x[0], x[1], x[2], y[0], y[1]);
        dlp_adj_apply(c, x[3], x[4], y[2], y[3], y[4]);
    }
};

// The filter on spectra in place, one thread per half-spectrum index; each
// thread reads its element of every input of a right-hand side before it
// writes any output.
template <int FK>
__global__ void k_filter(FilterArgs a)
{
    using Op = FilterOp<FK>;
    const int Mh = a.My / 2 + 1;
    const int64_t n = (int64_t)a.Mx * Mh;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const auto c = Op::coef(a, (int)(p / Mh), (int)(p % Mh));
    for (int r = 0; r < a.nrhs; ++r) {
        double2 x[Op::nin], y[Op::nout];
#pragma unroll
        for (int i = 0; i < Op::nin; ++i) x[i] = a.spec[(int64_t)Op::in(r, i) * a.sstride + p];
        Op::apply(c, x, y);
#pragma unroll
        for (int o = 0; o < Op::nout; ++o) a.spec[(int64_t)Op::out(r, o) * a.sstride + p] = y[o];
    }
}

// du of the points in bin order as two density columns of the gridding
// data: out[i] = du[2 perm[i]], out[n + i] = du[2 perm[i] + 1].
__global__ void k_gather_du_cols(const double *__restrict__ du, const int *__restrict__ perm, int n,
                                 double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    const double2 v = *(const double2 *)(du + 2 * (int64_t)o);
    out[i] = v.x;
    out[n + i] = v.y;
}

// ---- interpolation ---------------------------------------------------------
struct InterpArgs {
    const double *ptar;
    int64_t nt;
    const double *grids;   // [2*nr][Mx][My] real
    int64_t gstride;
    int Mx, My, P, nr;
    double Lx, Ly, h, c, w, ch2, scale;
    double *u;             // rhs r at u + r*ustride, interleaved, caller's order
    int64_t ustride;
};

template <int NR>
__global__ __launch_bounds__(256) void k_interp(InterpArgs a)
{
    __shared__ double E1[128];
    fill_e1(E1, a.P, a.ch2);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nt) return;
    int mx, my;
    double px, py;
    se2p_anchor(a.ptar[2 * t], a.ptar[2 * t + 1], a.Lx, a.Ly, a.h, a.P, &mx, &my, &px, &py);
    double bx, e3x, e3y;
    fgg_factors(px, py, a.c, a.w, a.h, &bx, &e3x, &e3y);
    double acc[NR][2];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r][0] = acc[r][1] = 0.0;
    int col = ((mx % a.Mx) + a.Mx) % a.Mx;
    const int row0 = ((my % a.My) + a.My) % a.My;
    double ex = bx;
    for (int i = 0; i <= a.P; ++i) {
        double ey = ex * E1[i];
        int row = row0;
        const int64_t cb = (int64_t)col * a.My;
        for (int j = 0; j <= a.P; ++j) {
            const double wgt = ey * E1[j];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                acc[r][0] += wgt * a.grids[(2 * r) * a.gstride + cb + row];
                acc[r][1] += wgt * a.grids[(2 * r + 1) * a.gstride + cb + row];
            }
            ey *= e3y;
            if (++row == a.My) row = 0;
        }
        ex *= e3x;
        if (++col == a.Mx) col = 0;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double *ur = a.u + (int64_t)r * a.ustride;
        ur[2 * t] = acc[r][0] * a.scale;
        ur[2 * t + 1] = acc[r][1] * a.scale;
    }
}


// ---- interpolation sums ----------------------------------------------------
// For a point and each of NR grids g_m: V_m = sum W g_m and the sums
// X_m = sum (i - P/2) W g_m, Y_m = sum (j - P/2) W g_m over the (P+1)^2 support
// nodes (i, j), W = ex E1[i] E1[j] e3x^i e3y^j the fast-Gaussian-gridding
// weight (mex_stokes_slp_kspace.cpp:264-329).  get(i, j, m) returns grid m at
// support node (i, j).
// The NR values of a support node stored as NR consecutive doubles (LDS);
// two grids are read as one 16-byte load.
template <int NR>
__device__ __forceinline__ void load_node(const double *p, double (&g)[NR])
{
    if constexpr (NR == 2) {
        const double2 v = *(const double2 *)p;
        g[0] = v.x;
        g[1] = v.y;
    } else {
#pragma unroll
        for (int m = 0; m < NR; ++m) g[m] = p[m];
    }
}

// get(i, j, g) fills g[0..NR) with the grid values at support node (i, j).
// DERIV false: V only (X, Y stay zero); V is computed the same way either way.
template <int NR, int PT, class GET, bool DERIV = true>
__device__ inline void interp_sums(const GET &get, int P, const double *E1, double bx, double e3x,
                                   double e3y, double (&V)[NR], double (&X)[NR], double (&Y)[NR])
{
#pragma unroll
    for (int m = 0; m < NR; ++m) V[m] = X[m] = Y[m] = 0.0;
    if constexpr (PT > 0) {
        // compile-time support: y weights and (j - P/2) y weights in registers,
        // one unrolled block per row
        constexpr double hq = 0.5 * PT;
        double wy[PT + 1], wj[PT + 1];
        double ey = 1.0;
#pragma unroll
        for (int j = 0; j <= PT; ++j) {
            wy[j] = ey * E1[j];
            wj[j] = (j - hq) * wy[j];
            ey *= e3y;
        }
        double ex = bx;
        for (int i = 0; i <= PT; ++i) {
            // even and odd nodes of the row in separate chains
            double ra[NR], rb[NR], ya[NR], yb[NR];
#pragma unroll
            for (int m = 0; m < NR; ++m) ra[m] = rb[m] = ya[m] = yb[m] = 0.0;
#pragma unroll
            for (int j = 0; j <= PT; j += 2) {
                // without the derivative sums: half a row of node loads in
                // flight at a time (120 VGPRs, four waves per SIMD, instead
                // of 146 and three)
                if (!DERIV && j == 2 * ((PT + 2) / 4)) __builtin_amdgcn_sched_barrier(0);
                double g[NR];
                get(i, j, g);
#pragma unroll
                for (int m = 0; m < NR; ++m) {
                    ra[m] = fma(wy[j], g[m], ra[m]);
                    if (DERIV) ya[m] = fma(wj[j], g[m], ya[m]);
                }
                if (j + 1 <= PT) {
                    get(i, j + 1, g);
#pragma unroll
                    for (int m = 0; m < NR; ++m) {
                        rb[m] = fma(wy[j + 1], g[m], rb[m]);
                        if (DERIV) yb[m] = fma(wj[j + 1], g[m], yb[m]);
                    }
                }
            }
            double r0[NR], ry[NR];
#pragma unroll
            for (int m = 0; m < NR; ++m) {
                r0[m] = ra[m] + rb[m];
                ry[m] = ya[m] + yb[m];
            }
            const double wx = ex * E1[i];
            const double ii = i - hq;
#pragma unroll
            for (int m = 0; m < NR; ++m) {
                const double v = wx * r0[m];
                V[m] += v;
                if (DERIV) {
                    X[m] = fma(ii, v, X[m]);
                    Y[m] = fma(wx, ry[m], Y[m]);
                }
            }
            ex *= e3x;
        }
        return;
    }
    const double hp = 0.5 * P;
    double ex = bx;
    for (int i = 0; i <= P; ++i) {
        double ey = ex * E1[i];
        double r0[NR], ry[NR];
#pragma unroll
        for (int m = 0; m < NR; ++m) r0[m] = ry[m] = 0.0;
        double jj = -hp;
        for (int j = 0; j <= P; ++j) {
            const double wgt = ey * E1[j];
            double gv[NR];
            get(i, j, gv);
#pragma unroll
            for (int m = 0; m < NR; ++m) {
                const double wg = wgt * gv[m];
                r0[m] += wg;
                ry[m] = fma(jj, wg, ry[m]);
            }
            ey *= e3y;
            jj += 1.0;
        }
        const double ii = i - hp;
#pragma unroll
        for (int m = 0; m < NR; ++m) {
            V[m] += r0[m];
            X[m] = fma(ii, r0[m], X[m]);
            Y[m] += ry[m];
        }
        ex *= e3x;
    }
}

// Gradients of the interpolated values, d(scale V_m)/dx and /dy, from the sums.
template <int NR>
__device__ inline void interp_grad(double c, double scale, double h, int P, double w, double px,
                                   double py, const double (&V)[NR], const double (&X)[NR],
                                   const double (&Y)[NR], double (&gx)[NR], double (&gy)[NR])
{
    const double k = -2.0 * c * scale;
    const double ox = h * (0.5 * P) - w - px, oy = h * (0.5 * P) - w - py;
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        gx[m] = k * (h * X[m] + ox * V[m]);
        gy[m] = k * (h * Y[m] + oy * V[m]);
    }
}

// dL/dptar from du and the gradients (gx1, gx2, gy1, gy2) of the two outputs.
__device__ __forceinline__ void target_grad(double a1, double a2, double gx0, double gx1,
                                            double gy0, double gy1, double *out)
{
    out[0] = a1 * gx0 + a2 * gx1;
    out[1] = a1 * gy0 + a2 * gy1;
}

// Tile variant: one workgroup per 16x16 block of target anchors.  The grid
// values the block's targets can touch, a (16+P)^2 region per grid, are staged
// in LDS once (as (G1, G2) pairs), then each thread interpolates one target at
// a time from LDS.  Targets come sorted by 8x8 anchor bin (the tile's targets
// are two runs of that order).
struct InterpTileArgs {
    const double *ptar;
    const int *perm;       // bin order -> caller's target index
    const int *boff;       // [nbx*nby+1] bin offsets of the targets
    int nbx, nby;
    const int *chunk_off;  // tile t owns chunks [chunk_off[t], chunk_off[t+1]) of 256 targets
    const int *block_tile; // workgroup -> tile
    int ntx, nty;
    const double *grids;   // [2*nr][Mx][My] real
    int64_t gstride;
    int Mx, My, P, nr;
    double Lx, Ly, h, c, w, ch2, scale;
    double *u;
    int64_t ustride;
    double *jac;           // DERIV: [4 nt] output gradients, caller order
    int64_t max_chunks;    // workgroups per right-hand side (k_interp_tile's grid)
};

// PT > 0: the support P = PT is a compile-time constant; each row of the
// support is then one unrolled block (its LDS loads issue together) and the
// separable y weights e3y^j E1[j] sit in registers, computed once per target.
// PT = 0: any P, weights formed node by node.
// DERIV (keeping the forward state): also the gradients of both outputs at
// every target, jac[4 t] = (du1/dx, du2/dx, du1/dy, du2/dy) in caller order,
// from which the backward call forms dL/dptar.
template <int PT, bool DERIV>
__device__ __forceinline__ void interp_tile_wg(const InterpTileArgs &a, const int b, const int rhs)
{
    extern __shared__ double lds[];
    const int P = PT > 0 ? PT : a.P;
    const int P1 = P + 1;
    const int R = kTile + P;              // region edge
    // row pitch R + 1: an odd number of 16-byte nodes, so that anchors a few
    // rows apart (points along a curve) fall on different LDS banks
    const int RP = R + 1;
    double2 *G = (double2 *)lds;          // [R][RP]
    double *E1 = (double *)(G + R * RP);  // [P1]
    const int ntiles = a.ntx * a.nty;
    if (b >= a.chunk_off[ntiles]) return;   // uniform over the workgroup
    const int tile = a.block_tile[b];
    const int chunk = b - a.chunk_off[tile];
    const int tx = tile / a.nty, ty = tile % a.nty;
    const int bx0 = 2 * tx, by0 = 2 * ty;
    const int nyb = min(2, a.nby - by0);
    int r0s = 0, r0e = 0, r1s = 0, r1e = 0;
    {
        const int k0 = bx0 * a.nby + by0;
        r0s = a.boff[k0];
        r0e = a.boff[k0 + nyb];
        if (bx0 + 1 < a.nbx) {
            const int k1 = (bx0 + 1) * a.nby + by0;
            r1s = a.boff[k1];
            r1e = a.boff[k1 + nyb];
        }
    }
    const int n0 = r0e - r0s, total = n0 + (r1e - r1s);
    const int v0 = chunk * 256, v1 = min(total, v0 + 256);
    fill_e1(E1, P, a.ch2);
    const int X0 = tx * kTile, Y0 = ty * kTile;
    // rhs selects the right-hand side
    for (int r = rhs; r < rhs + 1 && r < a.nr; ++r) {
        const double *g1 = a.grids + (int64_t)(2 * r) * a.gstride;
        const double *g2 = g1 + a.gstride;
        __syncthreads();   // previous RHS done with G
        for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
            const int i = e / R, j = e % R;
            int X = X0 + i, Y = Y0 + j;
            X %= a.Mx;
            Y %= a.My;
            const int64_t idx = (int64_t)X * a.My + Y;
            G[i * RP + j] = make_double2(g1[idx], g2[idx]);
        }
        __syncthreads();
        for (int v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
            const int k = v < n0 ? r0s + v : r1s + (v - n0);
            const int t = a.perm[k];
            int mx, my;
            double px, py;
            se2p_anchor(a.ptar[2 * t], a.ptar[2 * t + 1], a.Lx, a.Ly, a.h, P, &mx, &my, &px,
                        &py);
            double bx, e3x, e3y;
            fgg_factors(px, py, a.c, a.w, a.h, &bx, &e3x, &e3y);
            int li = ((mx % a.Mx) + a.Mx) % a.Mx - X0;
            int lj = ((my % a.My) + a.My) % a.My - Y0;
            double s1 = 0.0, s2 = 0.0;
            double ex = bx;
            const double2 *row = G + li * RP + lj;
            if constexpr (PT > 0) {
                const double *base = (const double *)row;
                auto get = [&](int i, int j, double (&g)[2]) { load_node<2>(base + (i * RP + j) * 2, g); };
                double V[2], X[2], Y[2];
                interp_sums<2, PT, decltype(get), DERIV>(get, PT, E1, bx, e3x, e3y, V, X, Y);
                s1 = V[0];
                s2 = V[1];
                if (DERIV) {
                    double gx[2], gy[2];
                    interp_grad<2>(a.c, a.scale, a.h, PT, a.w, px, py, V, X, Y, gx, gy);
                    *(double4 *)(a.jac + 4 * (int64_t)t) = make_double4(gx[0], gx[1], gy[0], gy[1]);
                }
            } else {
                for (int i = 0; i < P1; ++i) {
                    double ey = ex * E1[i];
                    for (int j = 0; j < P1; ++j) {
                        const double wgt = ey * E1[j];
                        const double2 gv = row[j];
                        s1 += wgt * gv.x;
                        s2 += wgt * gv.y;
                        ey *= e3y;
                    }
                    ex *= e3x;
                    row += RP;
                }
            }
            double *ur = a.u + (int64_t)r * a.ustride;
            ur[2 * t] = s1 * a.scale;
            ur[2 * t + 1] = s2 * a.scale;
        }
    }
}

// One-dimensional grid of max_chunks x right-hand sides, right-hand sides
// fastest (as k_spread_strip's groups).
template <int PT, bool DERIV>
__global__ __launch_bounds__(256) void k_interp_tile(InterpTileArgs a)
{
    const int nrhs = (int)(gridDim.x / (unsigned)a.max_chunks);
    const int b = (int)blockIdx.x / nrhs;
    interp_tile_wg<PT, DERIV>(a, b, (int)blockIdx.x - b * nrhs);
}

struct InterpPack {
    InterpTileArgs a[kPack];
    int off[kPack + 1];
    int n;
};
static_assert(sizeof(InterpPack) <= 4096, "kernel argument size");

// Many boxes in one launch, as k_spread_strip_boxes.
template <int PT>
__global__ __launch_bounds__(256) void k_interp_tile_boxes(InterpPack p)
{
    const int j = pack_slot(p.off, p.n, blockIdx.x);
    interp_tile_wg<PT, false>(p.a[j], blockIdx.x - p.off[j], blockIdx.y);
}

// ---- adjoint interpolation -------------------------------------------------
// For each point and each of NR grids g_m: V_m = sum W g_m and the sums
// X_m = sum (i - P/2) W g_m, Y_m = sum (j - P/2) W g_m over the (P+1)^2 support
// nodes (i, j).  The weight is W = exp(c (h i - w - px)^2 + c (h j - w - py)^2)
// (the factored form above, px the offset from the anchor node), so
//   dV_m/dx = -2c sum (h i - w - px) W g_m = -2c (h X_m + (h P/2 - w - px) V_m),
// the same for y.  The gradient is recomputed here, never stored.
enum AdjMode {
    ADJ_TARGET = 0,   // dL/dptar = scale * sum_m du_m grad V_m   (NR = 2)
    ADJ_SOURCE_SLP,   // dL/df = scale * V, dL/dpsrc = scale * sum_m f_m grad V_m
    ADJ_SOURCE_DLP    // V = dL/d(n1f1, n2f1 + n1f2, n2f2): dL/df, dL/dn, dL/dpsrc
};

struct InterpAdjArgs {
    const double *pts;     // points, caller order
    const int *perm;       // bin order -> caller's index (tile kernel)
    const int *boff;       // [nbx*nby+1] bin offsets of the points (tile kernel)
    int nbx, nby;
    const int *chunk_off;  // tile t owns chunks [chunk_off[t], chunk_off[t+1]) of 256 points
    const int *block_tile; // workgroup -> tile
    int ntx, nty;
    int64_t n;             // point count (direct kernel)
    const double *grids;   // [NR][Mx][My] real
    int64_t gstride;
    int Mx, My, P;
    double Lx, Ly, h, c, w, ch2, scale;
    const double *du;      // ADJ_TARGET: du at the points
    const double *f, *nv;  // ADJ_SOURCE_*: densities and normals at the points
    double *dpos, *df, *dn;   // outputs, caller order; may be null
};

template <int NR, int MODE>
__device__ inline void interp_adj_store(const InterpAdjArgs &a, int o, double px, double py,
                                        const double (&V)[NR], const double (&X)[NR],
                                        const double (&Y)[NR])
{
    double gx[NR], gy[NR];
    interp_grad<NR>(a.c, a.scale, a.h, a.P, a.w, px, py, V, X, Y, gx, gy);
    if (MODE == ADJ_TARGET) {
        target_grad(a.du[2 * o], a.du[2 * o + 1], gx[0], gx[NR > 1 ? 1 : 0], gy[0],
                    gy[NR > 1 ? 1 : 0], a.dpos + 2 * o);
    } else if (MODE == ADJ_SOURCE_SLP) {
        const double f1 = a.f[2 * o], f2 = a.f[2 * o + 1];
        if (a.dpos) {
            a.dpos[2 * o] = f1 * gx[0] + f2 * gx[1];
            a.dpos[2 * o + 1] = f1 * gy[0] + f2 * gy[1];
        }
        if (a.df) {
            a.df[2 * o] = a.scale * V[0];
            a.df[2 * o + 1] = a.scale * V[1];
        }
    } else {
        const double f1 = a.f[2 * o], f2 = a.f[2 * o + 1];
        const double n1 = a.nv[2 * o], n2 = a.nv[2 * o + 1];
        const double q1 = n1 * f1, q2 = n2 * f1 + n1 * f2, q3 = n2 * f2;
        const double m1 = a.scale * V[0], m2 = a.scale * V[NR > 1 ? 1 : 0],
                     m3 = a.scale * V[NR > 2 ? 2 : 0];
        if (a.dpos) {
            a.dpos[2 * o] = q1 * gx[0] + q2 * gx[NR > 1 ? 1 : 0] + q3 * gx[NR > 2 ? 2 : 0];
            a.dpos[2 * o + 1] = q1 * gy[0] + q2 * gy[NR > 1 ? 1 : 0] + q3 * gy[NR > 2 ? 2 : 0];
        }
        if (a.df) {
            a.df[2 * o] = m1 * n1 + m2 * n2;
            a.df[2 * o + 1] = m2 * n1 + m3 * n2;
        }
        if (a.dn) {
            a.dn[2 * o] = m1 * f1 + m2 * f2;
            a.dn[2 * o + 1] = m2 * f1 + m3 * f2;
        }
    }
}

__host__ __device__ inline size_t interp_adj_lds(int P, int NR)
{
    const size_t R = kTile + P;
    return sizeof(double) * (R * (R + 1) * NR + P + 1);
}

// Tile kernel: the (16+P)^2 region of the NR grids that a 16x16 block of
// anchors can touch is staged in LDS, as in k_interp_tile.
template <int NR, int MODE, int PT>
__global__ __launch_bounds__(256) void k_interp_adj_tile(InterpAdjArgs a)
{
    extern __shared__ double lds[];
    const int P = PT > 0 ? PT : a.P;
    const int P1 = P + 1;
    const int R = kTile + P;
    const int RP = R + 1;                      // row pitch (bank spread, as k_interp_tile)
    double *G = lds;                           // [R][RP][NR]
    double *E1 = G + (size_t)R * RP * NR;      // [P1]
    const int ntiles = a.ntx * a.nty;
    const int b = blockIdx.x;
    if (b >= a.chunk_off[ntiles]) return;   // uniform over the workgroup
    const int tile = a.block_tile[b];
    const int chunk = b - a.chunk_off[tile];
    const int tx = tile / a.nty, ty = tile % a.nty;
    const int bx0 = 2 * tx, by0 = 2 * ty;
    const int nyb = min(2, a.nby - by0);
    int r0s = 0, r0e = 0, r1s = 0, r1e = 0;
    {
        const int k0 = bx0 * a.nby + by0;
        r0s = a.boff[k0];
        r0e = a.boff[k0 + nyb];
        if (bx0 + 1 < a.nbx) {
            const int k1 = (bx0 + 1) * a.nby + by0;
            r1s = a.boff[k1];
            r1e = a.boff[k1 + nyb];
        }
    }
    const int n0 = r0e - r0s, total = n0 + (r1e - r1s);
    const int v0 = chunk * 256, v1 = min(total, v0 + 256);
    fill_e1(E1, P, a.ch2);
    const int X0 = tx * kTile, Y0 = ty * kTile;
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
        const int i = e / R, j = e % R;
        const int64_t idx = (int64_t)((X0 + i) % a.Mx) * a.My + (Y0 + j) % a.My;
#pragma unroll
        for (int m = 0; m < NR; ++m) G[(i * RP + j) * NR + m] = a.grids[m * a.gstride + idx];
    }
    __syncthreads();
    for (int v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const int k = v < n0 ? r0s + v : r1s + (v - n0);
        const int o = a.perm[k];
        int mx, my;
        double px, py;
        se2p_anchor(a.pts[2 * o], a.pts[2 * o + 1], a.Lx, a.Ly, a.h, P, &mx, &my, &px, &py);
        double bx, e3x, e3y;
        fgg_factors(px, py, a.c, a.w, a.h, &bx, &e3x, &e3y);
        const int li = ((mx % a.Mx) + a.Mx) % a.Mx - X0;
        const int lj = ((my % a.My) + a.My) % a.My - Y0;
        const double *base = G + (size_t)(li * RP + lj) * NR;
        auto get = [&](int i, int j, double (&g)[NR]) { load_node<NR>(base + (i * RP + j) * NR, g); };
        double V[NR], X[NR], Y[NR];
        interp_sums<NR, PT>(get, P, E1, bx, e3x, e3y, V, X, Y);
        interp_adj_store<NR, MODE>(a, o, px, py, V, X, Y);
    }
    (void)P1;
}

// Direct kernel (regions too large for LDS): one thread per point, grid
// values read from global memory with periodic wrap.
template <int NR, int MODE>
__global__ __launch_bounds__(256) void k_interp_adj_direct(InterpAdjArgs a)
{
    __shared__ double E1[128];   // P <= 126
    fill_e1(E1, a.P, a.ch2);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const int o = (int)t;
    int mx, my;
    double px, py;
    se2p_anchor(a.pts[2 * o], a.pts[2 * o + 1], a.Lx, a.Ly, a.h, a.P, &mx, &my, &px, &py);
    double bx, e3x, e3y;
    fgg_factors(px, py, a.c, a.w, a.h, &bx, &e3x, &e3y);
    const int col0 = ((mx % a.Mx) + a.Mx) % a.Mx;
    const int row0 = ((my % a.My) + a.My) % a.My;
    auto get = [&](int i, int j, double (&g)[NR]) {
        const int64_t idx = (int64_t)((col0 + i) % a.Mx) * a.My + (row0 + j) % a.My;
#pragma unroll
        for (int m = 0; m < NR; ++m) g[m] = a.grids[m * a.gstride + idx];
    };
    double V[NR], X[NR], Y[NR];
    interp_sums<NR, 0>(get, a.P, E1, bx, e3x, e3y, V, X, Y);
    interp_adj_store<NR, MODE>(a, o, px, py, V, X, Y);
}

// dL/dptar from the kept output gradients (k_interp_tile<PT, true>).
__global__ void k_target_grad(const double *__restrict__ jac, const double *__restrict__ du,
                              int64_t n, double *__restrict__ dptar)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double4 g = *(const double4 *)(jac + 4 * t);
    target_grad(du[2 * t], du[2 * t + 1], g.x, g.y, g.z, g.w, dptar + 2 * t);
}

// ---- rocFFT plans ------------------------------------------------------------
static int get_plans(se2p_ctx *ctx, int Mx, int My, int nfwd, int ninv, FftPlans **out)
{
    auto key = std::make_tuple(Mx, My, nfwd, ninv, ctx->plan_slot);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = &it->second;
        return 0;
    }
    if (int rc = refuse_if_frozen(ctx, "creating rocFFT plans for a new grid", 0)) return rc;
    FftPlans p;
    size_t lengths[2] = {(size_t)My, (size_t)Mx};   // fastest dimension first
    rocfft_status s;
    s = rocfft_plan_create(&p.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_double, 2, lengths, (size_t)nfwd, nullptr);
    if (s != rocfft_status_success) return fail(ctx, SE2P_ERR_FFT, "rocfft_plan_create (fwd) = %d", (int)s);
    s = rocfft_plan_create(&p.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_double, 2, lengths, (size_t)ninv, nullptr);
    if (s != rocfft_status_success) return fail(ctx, SE2P_ERR_FFT, "rocfft_plan_create (inv) = %d", (int)s);
    rocfft_plan_get_work_buffer_size(p.fwd, &p.fwd_work);
    rocfft_plan_get_work_buffer_size(p.inv, &p.inv_work);
    rocfft_execution_info_create(&p.fwd_info);
    rocfft_execution_info_create(&p.inv_info);
    size_t wmax = std::max(p.fwd_work, p.inv_work);
    if (wmax > 0) {
        void *wb = nullptr;
        SE2P_HIP(hipMalloc(&wb, wmax));
        p.work = wb;
        rocfft_execution_info_set_work_buffer(p.fwd_info, wb, p.fwd_work);
        rocfft_execution_info_set_work_buffer(p.inv_info, wb, p.inv_work);
    }
    ctx->plans[key] = p;
    *out = &ctx->plans[key];
    return 0;
}

int check_kspace_params(se2p_ctx *ctx, const KGeom &g)
{
    if (!(g.xi > 0) || !(g.eta > 0) || !(g.Lx > 0) || !(g.Ly > 0) || !std::isfinite(g.xi) ||
        !std::isfinite(g.eta) || !std::isfinite(g.Lx) || !std::isfinite(g.Ly))
        return fail(ctx, SE2P_ERR_PARAM, "xi, eta, Lx and Ly must be positive and finite");
    if (g.P < 0 || (g.P % 2) != 0 || g.P > 126)
        return fail(ctx, SE2P_ERR_P, "P must be even and in [0, 126] (got %d)", g.P);
    if (g.Mx <= 0 || g.My <= 0 || (g.Mx % 2) || (g.My % 2))
        return fail(ctx, SE2P_ERR_GRID, "Mx and My must be positive and even (got %d, %d)", g.Mx, g.My);
    if (g.Mx < g.P / 2 + 2 || g.My < g.P / 2 + 2)
        return fail(ctx, SE2P_ERR_GRID, "grid %dx%d too small for P = %d", g.Mx, g.My, g.P);
    const double hx = g.Lx / g.Mx, hy = g.Ly / g.My;
    if (std::fabs(hx - hy) > 1e-6 * hx)
        return fail(ctx, SE2P_ERR_GRID, "Lx/Mx (%.17g) != Ly/My (%.17g)", hx, hy);
    const double wexp = g.P * hx / 2;
    if (!(std::fabs(g.w - wexp) <= 1e-10 * std::max(wexp, 1e-300)) && !(g.P == 0 && g.w == 0))
        return fail(ctx, SE2P_ERR_W, "w (%.17g) must equal P*Lx/Mx/2 (%.17g)", g.w, wexp);
    return 0;
}

// ---- host building blocks ----------------------------------------------------
constexpr size_t kLdsMax = 65536;

// Gridding geometry of one call.
struct KLayout {
    int Mx, My, P;
    int nbx, nby, nbins;     // 8x8 anchor bins
    int ntx, nty, ntiles;    // 16x16 node tiles
    int64_t G, GH;           // real grid / half spectrum sizes
    double h, c, ch2, interp_scale;
    bool wrap;               // the support is wider than the grid
    size_t lds_itile;        // forward interpolation tile (2 grids)
    bool itile;
};

static KLayout make_layout(const KGeom &g)
{
    KLayout L;
    L.Mx = g.Mx;
    L.My = g.My;
    L.P = g.P;
    L.nbx = (L.Mx + kBin - 1) / kBin;
    L.nby = (L.My + kBin - 1) / kBin;
    L.nbins = L.nbx * L.nby;
    L.ntx = (L.Mx + kTile - 1) / kTile;
    L.nty = (L.My + kTile - 1) / kTile;
    L.ntiles = L.ntx * L.nty;
    L.G = (int64_t)L.Mx * L.My;
    L.GH = (int64_t)L.Mx * (L.My / 2 + 1);
    L.h = g.Lx / L.Mx;   // also used for y (mex_stokes_slp_kspace.cpp:45)
    L.c = -2 * g.xi * g.xi / g.eta;
    L.ch2 = L.c * L.h * L.h;
    double s = 4 * g.xi * g.xi / g.eta;   // :325-328
    s = s * s * L.h * L.h / SE2P_PI;
    L.interp_scale = s / (4 * SE2P_PI);
    L.wrap = (L.Mx <= L.P) || (L.My <= L.P);
    L.lds_itile = sizeof(double2) * (size_t)(kTile + L.P) * (kTile + L.P + 1) + sizeof(double) * (L.P + 1);
    L.itile = L.lds_itile <= kLdsMax;
    return L;
}

// One spreading pass (nsp points into ngrp groups of nq grids) and the
// interpolation at nip points that reads its result: kernel choice, work
// split and buffers.
struct GridPass {
    int nq, ngrp;
    bool strip;
    size_t lds_strip, lds_gen;
    int chunk;
    int64_t max_spread_chunks, max_interp_chunks;
    int *scnt, *soffc, *icnt, *ioffc, *rng_start, *rng_len, *smap, *imap;
    double *partial;
};

static void plan_pass(const KLayout &L, int nq, int ngrp, int64_t nsp, int64_t nip, int chunk,
                      GridPass *gp)
{
    const int P = L.P;
    gp->nq = nq;
    gp->ngrp = ngrp;
    gp->lds_strip = sizeof(double) * strip_lds_doubles(P, nq);
    gp->strip = (L.Mx > P) && (L.My > P + kPad) && gp->lds_strip <= kLdsMax;
    gp->lds_gen = sizeof(double) * ((size_t)2 * kBatch * (P + 1) + (size_t)kBatch * nq + (P + 1)) +
                  sizeof(int) * 2 * kBatch;
    // candidates per workgroup: the context's setting, else 4 wavefronts x 4 batches
    gp->chunk = chunk > 0 ? chunk : 4 * strip_batch(nq) * 4;
    const int64_t tps_x = std::min<int64_t>(L.ntx, (kTile + P + kBin - 1) / kTile + 1);
    const int64_t tps_y = std::min<int64_t>(L.nty, (kTile + P + kBin - 1) / kTile + 1);
    gp->max_spread_chunks = L.ntiles + (tps_x * tps_y * nsp + gp->chunk - 1) / gp->chunk;
    gp->max_interp_chunks = L.ntiles + (nip + 255) / 256;
}

static void carve_pass(Carve &cv, const KLayout &L, GridPass *gp)
{
    gp->scnt = cv.take<int>(L.ntiles + 1);
    gp->soffc = cv.take<int>(L.ntiles + 1);
    gp->icnt = cv.take<int>(L.ntiles + 1);
    gp->ioffc = cv.take<int>(L.ntiles + 1);
    gp->rng_start = cv.take<int>((size_t)L.ntiles * kTR);
    gp->rng_len = cv.take<int>((size_t)L.ntiles * kTR);
    gp->smap = cv.take<int>(gp->max_spread_chunks);
    gp->imap = cv.take<int>(gp->max_interp_chunks);
    gp->partial = gp->strip ? cv.take<double>((size_t)256 * gp->nq * gp->max_spread_chunks * gp->ngrp)
                            : nullptr;
}

static SpreadArgs spread_args(const KLayout &L, const GridPass &gp, const int *boff, double *grids)
{
    SpreadArgs sa{};
    sa.boff = boff;
    sa.nbx = L.nbx;
    sa.nby = L.nby;
    sa.Mx = L.Mx;
    sa.My = L.My;
    sa.P = L.P;
    sa.ch2 = L.ch2;
    sa.grids = grids;
    sa.gstride = L.G;
    sa.chunk_off = gp.soffc;
    sa.ntx = L.ntx;
    sa.nty = L.nty;
    sa.chunk = gp.chunk;
    sa.partial = gp.partial;
    sa.block_tile = gp.smap;
    sa.rng_start = gp.rng_start;
    sa.rng_len = gp.rng_len;
    sa.max_chunks = gp.max_spread_chunks;
    return sa;
}

static BinKeyFn bin_key_fn(const KLayout &L, const KGeom &g, const double *p)
{
    return BinKeyFn{p, g.Lx, g.Ly, L.h, L.P, L.Mx, L.My, L.nby};
}

// Sort n points by anchor bin.
static int bin_points(se2p_ctx *ctx, const KLayout &L, const KGeom &g, const double *p, int64_t n,
                      KeySort *ks)
{
    if (n > 0)
        hipLaunchKernelGGL(k_bin_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           bin_key_fn(L, g, p), (int)n, ks->key);
    return keysort_run(ctx, ks, n, L.nbins);
}

// Work split of a pass: candidate runs and chunk counts per tile for the
// spreading (bins sa.boff) and for the interpolation at the points binned in
// iboff.
static void tile_maps(hipStream_t st, const KLayout &L, const GridPass &gp, const SpreadArgs &sa,
                      const int *iboff)
{
    hipLaunchKernelGGL(k_tile_work, dim3(L.ntiles), dim3(64), 0, st, sa, iboff, gp.scnt, gp.icnt,
                       gp.rng_start, gp.rng_len);
    plan_chunks2(st, gp.scnt, L.ntiles, gp.soffc, gp.smap, gp.icnt, L.ntiles, gp.ioffc, gp.imap);
}

static int spread(se2p_ctx *ctx, const KLayout &L, const GridPass &gp, const SpreadArgs &sa)
{
    hipStream_t st = ctx->stream;
    const dim3 tiles2d(L.ntx, L.nty, gp.ngrp);
    const bool wide = L.Mx >= L.P + 16 && L.My >= L.P + 16;
#define SE2P_SPREAD(NQV)                                                                           \
    if (gp.strip && wide) {                                                                        \
        hipLaunchKernelGGL((k_spread_strip<NQV, true>),                                            \
                           dim3((unsigned)(gp.max_spread_chunks * gp.ngrp)), dim3(256), gp.lds_strip, \
                           st, sa);                                                                \
        hipLaunchKernelGGL((k_spread_combine<NQV>), dim3(L.ntiles, gp.ngrp), dim3(64), 0, st, sa); \
    } else if (gp.strip) {                                                                         \
        hipLaunchKernelGGL((k_spread_strip<NQV, false>),                                           \
                           dim3((unsigned)(gp.max_spread_chunks * gp.ngrp)), dim3(256), gp.lds_strip, \
                           st, sa);                                                                \
        hipLaunchKernelGGL((k_spread_combine<NQV>), dim3(L.ntiles, gp.ngrp), dim3(64), 0, st, sa); \
    } else if (L.wrap) {                                                                           \
        hipLaunchKernelGGL((k_spread<NQV, true>), tiles2d, dim3(256), gp.lds_gen, st, sa);        \
    } else {                                                                                       \
        hipLaunchKernelGGL((k_spread<NQV, false>), tiles2d, dim3(256), gp.lds_gen, st, sa);       \
    }
    switch (gp.nq) {
    case 2: SE2P_SPREAD(2); break;
    case 3: SE2P_SPREAD(3); break;
    case 4: SE2P_SPREAD(4); break;
    case 5: SE2P_SPREAD(5); break;
    case 8: SE2P_SPREAD(8); break;
    case 12: SE2P_SPREAD(12); break;
    default: return fail(ctx, SE2P_ERR_ARG, "unsupported grid count %d", gp.nq);
    }
#undef SE2P_SPREAD
    SE2P_HIP(hipGetLastError());
    return 0;
}

// Forward transforms, filter, inverse transforms; grids in, filtered real
// grids out (spec is scratch).
static int transform_filter(se2p_ctx *ctx, const KLayout &L, const KGeom &g, FftPlans *plans,
                            FilterKind fk, int nrhs, double *grids, double2 *spec)
{
    hipStream_t st = ctx->stream;
    stage_begin(ctx, ST_FFT_FWD);
    rocfft_execution_info_set_stream(plans->fwd_info, st);
    {
        void *in[1] = {grids};
        void *out[1] = {spec};
        rocfft_status s = rocfft_execute(plans->fwd, in, out, plans->fwd_info);
        if (s != rocfft_status_success) return fail(ctx, SE2P_ERR_FFT, "rocfft_execute (fwd) = %d", (int)s);
    }
    stage_end(ctx, ST_FFT_FWD);

    stage_begin(ctx, ST_FILTER);
    FilterArgs fa;
    fa.spec = spec;
    fa.sstride = L.GH;
    fa.Mx = L.Mx;
    fa.My = L.My;
    fa.nrhs = nrhs;
    fa.Lx = g.Lx;
    fa.Ly = g.Ly;
    fa.xi = g.xi;
    fa.eta = g.eta;
    fa.scale = 1.0 / ((double)L.Mx * (double)L.My);
    const unsigned fblocks = (unsigned)((L.GH + 255) / 256);
    if (fk == FILTER_SLP) hipLaunchKernelGGL(k_filter<FILTER_SLP>, dim3(fblocks), dim3(256), 0, st, fa);
    else if (fk == FILTER_DLP) hipLaunchKernelGGL(k_filter<FILTER_DLP>, dim3(fblocks), dim3(256), 0, st, fa);
    else if (fk == FILTER_DLP_ADJ) hipLaunchKernelGGL(k_filter<FILTER_DLP_ADJ>, dim3(fblocks), dim3(256), 0, st, fa);
    else if (fk == FILTER_BOTH_SLP) hipLaunchKernelGGL(k_filter<FILTER_BOTH_SLP>, dim3(fblocks), dim3(256), 0, st, fa);
    else hipLaunchKernelGGL(k_filter<FILTER_BOTH_DLP>, dim3(fblocks), dim3(256), 0, st, fa);
    SE2P_HIP(hipGetLastError());
    stage_end(ctx, ST_FILTER);

    stage_begin(ctx, ST_FFT_INV);
    rocfft_execution_info_set_stream(plans->inv_info, st);
    {
        void *in[1] = {spec};
        void *out[1] = {grids};
        rocfft_status s = rocfft_execute(plans->inv, in, out, plans->inv_info);
        if (s != rocfft_status_success) return fail(ctx, SE2P_ERR_FFT, "rocfft_execute (inv) = %d", (int)s);
    }
    stage_end(ctx, ST_FFT_INV);
    return 0;
}

// State a forward call keeps for its backward call (se2p_set_keep): one
// single-density, targets-are-sources call, identified by its arguments.
// The bins, Gaussian factors, work plans and the filtered grids live in a
// buffer of their own.
struct KeptK {
    char *buf = nullptr;
    size_t bytes = 0;
    bool valid = false;
    int kind = 0;
    KGeom g{};
    const double *f = nullptr, *nv = nullptr;
    KeySort kss{};
    GridPass gp{};
    double *grids = nullptr;   // the two filtered grids
    double *jac = nullptr;     // output gradients at the targets, or null
    bool match(int k, const KGeom &q, const double *ff, const double *nn) const
    {
        return valid && k == kind && q.psrc == g.psrc && q.ptar == g.psrc && q.ns == g.ns &&
               q.nt == g.ns && ff == f && (k == 0 || nn == nv) && q.xi == g.xi && q.eta == g.eta &&
               q.Lx == g.Lx && q.Ly == g.Ly && q.w == g.w && q.Mx == g.Mx && q.My == g.My &&
               q.P == g.P;
    }
};

void free_kept_k(se2p_ctx *ctx)
{
    if (!ctx->kept_k) return;
    if (ctx->kept_k->buf) (void)hipFree(ctx->kept_k->buf);
    delete ctx->kept_k;
    ctx->kept_k = nullptr;
}

// ---- forward -------------------------------------------------------------------
// One Fourier-space sum's plan and buffers: a single-box call, or one box of a
// multi-box call.
struct KRun {
    int kind = 0, nrhs = 1, ncomp = 2, RB = 1, ngrp = 1, nrp = 1;
    KGeom g{};
    const double *f = nullptr, *nv = nullptr;
    int64_t dstride = 0, ustride = 0;
    double *u = nullptr;
    bool same_points = false;   // targets are the source array: one set of bins
    bool maps = false;          // tile work maps are used (strip spreading or tile interpolation)
    bool gathered = false;      // the fused sort gathered the gridding data
    bool want_jac = false;      // keep the output gradients for a backward call
    bool shared_sort = false;   // bins and gridding data come from a sort of many boxes
    KLayout L{};
    GridPass gp{};
    KeySort kss{}, kst{};
    int2 *anchor = nullptr;
    double *fgg = nullptr, *q = nullptr, *grids = nullptr, *jac = nullptr;
    double2 *spec = nullptr;
    SpreadArgs sa{};
    FftPlans *plans = nullptr;

    const KeySort &tbins() const { return same_points ? kss : kst; }

    void layout(Carve &cv)
    {
        const int64_t ns = g.ns, nt = g.nt;
        if (want_jac) jac = cv.take<double>(4 * (size_t)nt);
        if (!shared_sort) {
            char *p = cv.take<char>(keysort_bytes(ns, L.nbins));
            if (p) keysort_carve(p, ns, L.nbins, &kss);
            if (!same_points) {
                p = cv.take<char>(keysort_bytes(nt, L.nbins));
                if (p) keysort_carve(p, nt, L.nbins, &kst);
            }
            anchor = cv.take<int2>(ns);
            fgg = cv.take<double>(3 * (size_t)ns);
            q = cv.take<double>((size_t)ncomp * nrp * ns);
        }
        grids = cv.take<double>((size_t)ncomp * nrp * L.G);
        spec = cv.take<double2>((size_t)ncomp * nrp * L.GH);
        carve_pass(cv, L, &gp);
    }
};

static void krun_init(KRun &r, int kind, const KGeom &g, const double *f, const double *nv,
                      int nrhs, int64_t dstride, int64_t ustride, double *u, int chunk)
{
    r.kind = kind;
    r.g = g;
    r.f = f;
    r.nv = nv;
    r.nrhs = nrhs;
    r.dstride = dstride;
    r.ustride = ustride;
    r.u = u;
    r.L = make_layout(g);
    r.ncomp = kind == 0 ? 2 : 3;
    // right-hand sides per spreading group (grids per group nq = ncomp*RB);
    // the last group is padded with zero densities
    r.RB = nrhs >= 2 ? 4 : 1;
    r.ngrp = (nrhs + r.RB - 1) / r.RB;
    r.nrp = r.ngrp * r.RB;
    // targets are binned separately unless they are the source array itself
    r.same_points = (g.ptar == g.psrc) && (g.nt == g.ns);
    plan_pass(r.L, r.ncomp * r.RB, r.ngrp, g.ns, g.nt, chunk, &r.gp);
    r.maps = r.gp.strip || r.L.itile;
}

// Bins of the sources (and of the targets unless they are the sources), the
// sources' gridding data and the work maps.  The fused sort (se2p_fsort.h)
// also gathers the sources' gridding data and, beside the sort that
// completes the target bins, builds the tile work lists and chunk plans;
// otherwise separate launches do.
static int krun_sort(se2p_ctx *ctx, KRun &r)
{
    hipStream_t st = ctx->stream;
    const KLayout &L = r.L;
    const KGeom &g = r.g;
    GridPass &gp = r.gp;
    const int64_t ns = g.ns, nt = g.nt;
    const KeySort &tbins = r.tbins();
    SpreadArgs &sa = r.sa;
    sa = spread_args(L, gp, r.kss.off, r.grids);
    sa.s.anchor = r.anchor;
    sa.s.fgg = r.fgg;
    sa.s.q = r.q;
    sa.s.ns = ns;
    const TileWorkExtra tw{sa, tbins.off, gp.scnt, gp.icnt, gp.rng_start, gp.rng_len,
                           r.maps ? L.ntiles : 0};
    const FsPlanExtra pl{PlanArgs{gp.scnt, nullptr, 1, 1, L.ntiles, gp.soffc, gp.smap},
                         PlanArgs{gp.icnt, nullptr, 1, 1, L.ntiles, gp.ioffc, gp.imap},
                         r.maps ? 2 : 0};
    const TileWorkExtra no_tw{sa, tbins.off, nullptr, nullptr, nullptr, nullptr, 0};
    const FsPlanExtra no_pl{pl.p0, pl.p1, 0};
    bool mapped = false;
    r.gathered = false;
    if (fsort_ok(ctx, ns, L.nbins)) {
        const BinKeyFn kf = bin_key_fn(L, g, g.psrc);
        int rc;
        if (r.kind == 0)
            rc = fsort_run(ctx, &r.kss, ns, L.nbins, kf,
                           GridSrcGather<0>{g.psrc, r.f, r.nv, (int)ns, r.nrhs, r.nrp, r.dstride, g.Lx,
                                            g.Ly, L.h, L.P, L.Mx, L.My, L.c, g.w, r.anchor, r.fgg, r.q},
                           r.same_points ? tw : no_tw, r.same_points ? pl : no_pl);
        else
            rc = fsort_run(ctx, &r.kss, ns, L.nbins, kf,
                           GridSrcGather<1>{g.psrc, r.f, r.nv, (int)ns, r.nrhs, r.nrp, r.dstride, g.Lx,
                                            g.Ly, L.h, L.P, L.Mx, L.My, L.c, g.w, r.anchor, r.fgg, r.q},
                           r.same_points ? tw : no_tw, r.same_points ? pl : no_pl);
        if (rc) return rc;
        r.gathered = true;
        mapped = r.same_points;
    } else if (int rc = bin_points(ctx, L, g, g.psrc, ns, &r.kss)) {
        return rc;
    }
    if (!r.same_points) {
        if (fsort_ok(ctx, nt, L.nbins)) {
            if (int rc = fsort_run(ctx, &r.kst, nt, L.nbins, bin_key_fn(L, g, g.ptar), FsNoPayload{},
                                   tw, pl))
                return rc;
            mapped = true;
        } else if (int rc = bin_points(ctx, L, g, g.ptar, nt, &r.kst)) {
            return rc;
        }
    }
    if (r.maps && !mapped) tile_maps(st, L, gp, sa, tbins.off);
    return 0;
}

// The sources' gridding data in bin order, unless the fused sort wrote it.
static void krun_gather(se2p_ctx *ctx, const KRun &r)
{
    if (r.gathered) return;
    const KLayout &L = r.L;
    const KGeom &g = r.g;
    const unsigned sblocks = (unsigned)((g.ns + 255) / 256);
    if (r.kind == 0)
        hipLaunchKernelGGL(k_gather_grid_src<0>, dim3(sblocks), dim3(256), 0, ctx->stream, g.psrc,
                           r.f, r.nv, r.kss.perm, (int)g.ns, r.nrhs, r.nrp, r.dstride, g.Lx, g.Ly,
                           L.h, L.P, L.Mx, L.My, L.c, g.w, r.anchor, r.fgg, r.q);
    else
        hipLaunchKernelGGL(k_gather_grid_src<1>, dim3(sblocks), dim3(256), 0, ctx->stream, g.psrc,
                           r.f, r.nv, r.kss.perm, (int)g.ns, r.nrhs, r.nrp, r.dstride, g.Lx, g.Ly,
                           L.h, L.P, L.Mx, L.My, L.c, g.w, r.anchor, r.fgg, r.q);
}

static InterpTileArgs interp_tile_args(const KRun &r)
{
    const KLayout &L = r.L;
    const KeySort &tbins = r.tbins();
    InterpTileArgs ia;
    ia.ptar = r.g.ptar;
    ia.perm = tbins.perm;
    ia.boff = tbins.off;
    ia.nbx = L.nbx;
    ia.nby = L.nby;
    ia.chunk_off = r.gp.ioffc;
    ia.block_tile = r.gp.imap;
    ia.ntx = L.ntx;
    ia.nty = L.nty;
    ia.grids = r.grids;
    ia.gstride = L.G;
    ia.Mx = L.Mx;
    ia.My = L.My;
    ia.P = L.P;
    ia.nr = r.nrhs;
    ia.Lx = r.g.Lx;
    ia.Ly = r.g.Ly;
    ia.h = L.h;
    ia.c = L.c;
    ia.w = r.g.w;
    ia.ch2 = L.ch2;
    ia.scale = L.interp_scale;
    ia.u = r.u;
    ia.ustride = r.ustride;
    ia.jac = r.jac;
    ia.max_chunks = r.gp.max_interp_chunks;
    return ia;
}

// Interpolation of the filtered grids at the targets.
static int krun_interp(se2p_ctx *ctx, const KRun &r)
{
    hipStream_t st = ctx->stream;
    const KLayout &L = r.L;
    const GridPass &gp = r.gp;
    const int nrhs = r.nrhs;
    if (L.itile) {
        const InterpTileArgs ia = interp_tile_args(r);
        if (L.P == 24 && r.jac)
            hipLaunchKernelGGL((k_interp_tile<24, true>), dim3((unsigned)(gp.max_interp_chunks * nrhs)),
                               dim3(256), L.lds_itile, st, ia);
        else if (L.P == 24)
            hipLaunchKernelGGL((k_interp_tile<24, false>), dim3((unsigned)(gp.max_interp_chunks * nrhs)),
                               dim3(256), L.lds_itile, st, ia);
        else
            hipLaunchKernelGGL((k_interp_tile<0, false>), dim3((unsigned)(gp.max_interp_chunks * nrhs)),
                               dim3(256), L.lds_itile, st, ia);
    } else {
        for (int rr = 0; rr < nrhs; ++rr) {
            InterpArgs ia;
            ia.ptar = r.g.ptar;
            ia.nt = r.g.nt;
            ia.grids = r.grids + (int64_t)(2 * rr) * L.G;
            ia.gstride = L.G;
            ia.Mx = L.Mx;
            ia.My = L.My;
            ia.P = L.P;
            ia.nr = 1;
            ia.Lx = r.g.Lx;
            ia.Ly = r.g.Ly;
            ia.h = L.h;
            ia.c = L.c;
            ia.w = r.g.w;
            ia.ch2 = L.ch2;
            ia.scale = L.interp_scale;
            ia.u = r.u + (int64_t)rr * r.ustride;
            ia.ustride = r.ustride;
            hipLaunchKernelGGL(k_interp<1>, dim3((unsigned)((r.g.nt + 255) / 256)), dim3(256), 0, st, ia);
        }
    }
    SE2P_HIP(hipGetLastError());
    return 0;
}

int kspace_sum(se2p_ctx *ctx, int kind, const KGeom &g, const double *f, const double *nv,
               int nrhs, int64_t dstride, int64_t ustride, double *u)
{
    hipStream_t st = ctx->stream;
    if (g.nt == 0 || nrhs == 0) return 0;
    if (g.ns == 0) {
        // no sources: every grid is zero, so is the result
        for (int r = 0; r < nrhs; ++r) launch_zero(st, u + (int64_t)r * ustride, 2 * g.nt);
        SE2P_HIP(hipGetLastError());
        return 0;
    }
    KRun r;
    krun_init(r, kind, g, f, nv, nrhs, dstride, ustride, u, ctx->spread_chunk);
    if (int rc = get_plans(ctx, r.L.Mx, r.L.My, r.ncomp * r.nrp, 2 * r.nrp, &r.plans)) return rc;

    // a call whose backward may follow keeps its state in a buffer of its own
    const bool keep = ctx->keep && !ctx->host_ptrs && r.same_points && nrhs == 1;
    r.want_jac = keep && r.L.itile && r.L.P == 24;
    Carve sizing;
    r.layout(sizing);
    Carve cv;
    if (keep) {
        if (!ctx->kept_k) ctx->kept_k = new KeptK;
        KeptK &kk = *ctx->kept_k;
        kk.valid = false;
        if (sizing.off > kk.bytes) {
            if (int rc = refuse_if_frozen(ctx, "growing the kept Fourier-space state", sizing.off))
                return rc;
            if (kk.buf) {
                SE2P_HIP(hipStreamSynchronize(st));
                SE2P_HIP(hipFree(kk.buf));
                kk.buf = nullptr;
                kk.bytes = 0;
            }
            SE2P_HIP(hipMalloc((void **)&kk.buf, sizing.off));
            kk.bytes = sizing.off;
        }
        cv.base = kk.buf;
    } else {
        if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
        cv.base = ctx->ws;
    }
    r.layout(cv);

    stage_begin(ctx, ST_SORT);
    if (int rc = krun_sort(ctx, r)) return rc;
    stage_end(ctx, ST_SORT);

    stage_begin(ctx, ST_SPREAD);
    krun_gather(ctx, r);
    if (int rc = spread(ctx, r.L, r.gp, r.sa)) return rc;
    stage_end(ctx, ST_SPREAD);

    if (int rc = transform_filter(ctx, r.L, g, r.plans, kind == 0 ? FILTER_SLP : FILTER_DLP, r.nrp,
                                  r.grids, r.spec))
        return rc;
    if (keep) {
        KeptK &kk = *ctx->kept_k;
        kk.kind = kind;
        kk.g = g;
        kk.f = f;
        kk.nv = nv;
        kk.kss = r.kss;
        kk.gp = r.gp;
        kk.grids = r.grids;
        kk.jac = r.jac;
        kk.valid = true;
    }

    stage_begin(ctx, ST_INTERP);
    if (int rc = krun_interp(ctx, r)) return rc;
    stage_end(ctx, ST_INTERP);
    return 0;
}

// ---- many boxes: one bin sort ----------------------------------------------------
// When every box's targets are its sources, the bins of all boxes come from
// one stable sort of global keys bin_base[box] + bin, so each box's points
// keep the order, and its gridding data the values, of a single-box call.
// The gridding data are gathered into arrays over all points (sorted
// positions, component stride n); a box's bin offsets are its slice of the
// global offsets (global positions), and its permutation entries are
// rewritten to box-local point indices.
struct MbBox {
    const double *p, *f, *nv;
    double Lx, Ly, h, c, w;
    int P, Mx, My, nby;
    int s_base, bin_base;
};
struct MbBoxBatch {
    MbBox b[kPack];
    int first, count;
};
static_assert(sizeof(MbBoxBatch) <= 4096, "kernel argument size");

__global__ void k_write_mb(MbBoxBatch bb, MbBox *__restrict__ tab, int *__restrict__ s_off, int nbox,
                           int n)
{
    const int i = threadIdx.x;
    if (i < bb.count) {
        tab[bb.first + i] = bb.b[i];
        s_off[bb.first + i] = bb.b[i].s_base;
    }
    if (i == 0 && bb.first + bb.count == nbox) s_off[nbox] = n;
}

// Box of global point g: the largest b with s_off[b] <= g (every box has
// points, so the offsets ascend strictly).
__device__ __forceinline__ int mb_box_of(const int *s_off, int nbox, int g)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_off[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct MbKeyFn {
    const MbBox *tab;
    const int *s_off;
    int nbox;
    __device__ int operator()(int g) const
    {
        const int b = mb_box_of(s_off, nbox, g);
        const MbBox &B = tab[b];
        const BinKeyFn kf{B.p, B.Lx, B.Ly, B.h, B.P, B.Mx, B.My, B.nby};
        return B.bin_base + kf(g - s_off[b]);
    }
};

__global__ void k_mb_keys(MbKeyFn kf, int n, int *__restrict__ keys)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) keys[g] = kf(g);
}

// Gridding data at sorted position i (as a single-box gather) and the
// permutation entry made box-local.
template <int KIND>
__global__ void k_mb_gather(const MbBox *__restrict__ tab, const int *__restrict__ s_off, int nbox,
                            int n, int *__restrict__ perm, int2 *__restrict__ anchor,
                            double *__restrict__ fgg, double *__restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = perm[i];
    const int b = mb_box_of(s_off, nbox, g);
    const MbBox &B = tab[b];
    const int o = g - s_off[b];
    const int64_t nsb = s_off[b + 1] - s_off[b];
    gather_grid_source<KIND>(B.p, B.f, B.nv, o, i, n, 1, 1, 2 * nsb, B.Lx, B.Ly, B.h, B.P, B.Mx,
                             B.My, B.c, B.w, anchor, fgg, q);
    perm[i] = o;
}

// Tile work lists of up to kPack boxes (one tile per workgroup) and their
// chunk plans (one plan per workgroup).
struct TileWorkPack {
    SpreadArgs a[kPack];
    int *scnt[kPack], *icnt[kPack], *rs[kPack], *rl[kPack];
    int off[kPack + 1];
    int n;
};
static_assert(sizeof(TileWorkPack) <= 4096, "kernel argument size");

__global__ __launch_bounds__(64) void k_tile_work_boxes(TileWorkPack p)
{
    const int j = pack_slot(p.off, p.n, blockIdx.x);
    tile_work(p.a[j], p.a[j].boff, blockIdx.x - p.off[j], threadIdx.x, p.scnt[j], p.icnt[j],
              p.rs[j], p.rl[j]);
}

struct PlanPack {
    PlanArgs p[2 * kPack];
    int n;
};
static_assert(sizeof(PlanPack) <= 4096, "kernel argument size");

__global__ __launch_bounds__(256) void k_plan_boxes(PlanPack pk)
{
    __shared__ int wsum[4];
    if ((int)blockIdx.x < pk.n) plan_body<4>(pk.p[blockIdx.x], wsum);
}

// The one-sort path's buffers over all points.
struct MbSort {
    int n = 0, nkeys = 0;
    MbBox *tab = nullptr;
    int *s_off = nullptr;
    KeySort ks{};
    int2 *anchor = nullptr;
    double *fgg = nullptr, *q = nullptr;
    void layout(Carve &cv, int nbox, int ncomp)
    {
        tab = cv.take<MbBox>((size_t)nbox);
        s_off = cv.take<int>((size_t)nbox + 1);
        char *p = cv.take<char>(keysort_bytes(n, nkeys));
        if (p) keysort_carve(p, n, nkeys, &ks);
        anchor = cv.take<int2>((size_t)n);
        fgg = cv.take<double>(3 * (size_t)n);
        q = cv.take<double>((size_t)ncomp * n);
    }
};

// Sorts and gathers the points of every run and builds their work maps; the
// runs then spread and interpolate as after krun_sort.
static int mb_sort(se2p_ctx *ctx, int kind, std::vector<KRun> &runs, MbSort &m)
{
    hipStream_t st = ctx->stream;
    const int nrun = (int)runs.size();
    for (int b0 = 0; b0 < nrun; b0 += kPack) {
        MbBoxBatch bb{};
        bb.first = b0;
        bb.count = std::min(kPack, nrun - b0);
        int s_base = 0, bin_base = 0;
        for (int k = 0; k < b0; ++k) {
            s_base += (int)runs[(size_t)k].g.ns;
            bin_base += runs[(size_t)k].L.nbins;
        }
        for (int i = 0; i < bb.count; ++i) {
            const KRun &r = runs[(size_t)(b0 + i)];
            bb.b[i] = MbBox{r.g.psrc, r.f, r.nv, r.g.Lx, r.g.Ly, r.L.h, r.L.c, r.g.w, r.L.P, r.L.Mx,
                            r.L.My, r.L.nby, s_base, bin_base};
            s_base += (int)r.g.ns;
            bin_base += r.L.nbins;
        }
        hipLaunchKernelGGL(k_write_mb, dim3(1), dim3(kPack), 0, st, bb, m.tab, m.s_off, nrun, m.n);
    }
    const MbKeyFn kf{m.tab, m.s_off, nrun};
    if (fsort_ok(ctx, m.n, m.nkeys)) {
        if (int rc = fsort_run(ctx, &m.ks, m.n, m.nkeys, kf, FsNoPayload{}, FsNoExtra{}, FsNoExtra{}))
            return rc;
    } else {
        hipLaunchKernelGGL(k_mb_keys, dim3((unsigned)((m.n + 255) / 256)), dim3(256), 0, st, kf, m.n,
                           m.ks.key);
        if (int rc = keysort_run(ctx, &m.ks, m.n, m.nkeys)) return rc;
    }
    const unsigned blocks = (unsigned)((m.n + 255) / 256);
    if (kind == 0)
        hipLaunchKernelGGL(k_mb_gather<0>, dim3(blocks), dim3(256), 0, st, m.tab, m.s_off, nrun, m.n,
                           m.ks.perm, m.anchor, m.fgg, m.q);
    else
        hipLaunchKernelGGL(k_mb_gather<1>, dim3(blocks), dim3(256), 0, st, m.tab, m.s_off, nrun, m.n,
                           m.ks.perm, m.anchor, m.fgg, m.q);
    // each run's views: global positions throughout, bins sliced per box
    std::vector<int> mapped;
    int bin_base = 0;
    for (int k = 0; k < nrun; ++k) {
        KRun &r = runs[(size_t)k];
        r.kss.perm = m.ks.perm;
        r.kss.off = m.ks.off + bin_base;
        bin_base += r.L.nbins;
        r.anchor = m.anchor;
        r.fgg = m.fgg;
        r.q = m.q;
        r.gathered = true;
        r.sa = spread_args(r.L, r.gp, r.kss.off, r.grids);
        r.sa.s.anchor = m.anchor;
        r.sa.s.fgg = m.fgg;
        r.sa.s.q = m.q;
        r.sa.s.ns = m.n;
        if (r.maps) mapped.push_back(k);
    }
    for (size_t s0 = 0; s0 < mapped.size(); s0 += kPack) {
        const int cnt = (int)std::min<size_t>(kPack, mapped.size() - s0);
        TileWorkPack tp{};
        PlanPack pp{};
        for (int j = 0; j < cnt; ++j) {
            KRun &r = runs[(size_t)mapped[s0 + (size_t)j]];
            tp.a[j] = r.sa;
            tp.scnt[j] = r.gp.scnt;
            tp.icnt[j] = r.gp.icnt;
            tp.rs[j] = r.gp.rng_start;
            tp.rl[j] = r.gp.rng_len;
            tp.off[j + 1] = tp.off[j] + r.L.ntiles;
            pp.p[2 * j] = PlanArgs{r.gp.scnt, nullptr, 1, 1, r.L.ntiles, r.gp.soffc, r.gp.smap};
            pp.p[2 * j + 1] = PlanArgs{r.gp.icnt, nullptr, 1, 1, r.L.ntiles, r.gp.ioffc, r.gp.imap};
        }
        tp.n = cnt;
        pp.n = 2 * cnt;
        hipLaunchKernelGGL(k_tile_work_boxes, dim3((unsigned)tp.off[cnt]), dim3(64), 0, st, tp);
        hipLaunchKernelGGL(k_plan_boxes, dim3((unsigned)pp.n), dim3(256), 0, st, pp);
    }
    SE2P_HIP(hipGetLastError());
    return 0;
}

// ---- many boxes ------------------------------------------------------------------
// Strip spreading of many boxes, kPack boxes per launch (each box's workgroups
// run exactly as in its own launch, so the grids are bitwise those of
// single-box calls); boxes the strip kernel does not take run one by one.
static int spread_boxes(se2p_ctx *ctx, std::vector<KRun> &runs)
{
    hipStream_t st = ctx->stream;
    std::vector<int> pk;
    for (size_t k = 0; k < runs.size(); ++k) {
        const KRun &r = runs[k];
        const bool wide = r.L.Mx >= r.L.P + 16 && r.L.My >= r.L.P + 16;
        if (r.gp.strip && wide && r.gp.ngrp == 1 && r.gp.nq == runs[0].gp.nq) pk.push_back((int)k);
        else if (int rc = spread(ctx, r.L, r.gp, r.sa)) return rc;
    }
    for (size_t s0 = 0; s0 < pk.size(); s0 += kPack) {
        const int n = (int)std::min<size_t>(kPack, pk.size() - s0);
        SpreadPack sp{}, cp{};
        size_t lds = 0;
        for (int j = 0; j < n; ++j) {
            const KRun &r = runs[(size_t)pk[s0 + (size_t)j]];
            sp.a[j] = cp.a[j] = r.sa;
            sp.off[j + 1] = sp.off[j] + (int)r.gp.max_spread_chunks;
            cp.off[j + 1] = cp.off[j] + r.L.ntiles;
            lds = std::max(lds, r.gp.lds_strip);
        }
        sp.n = cp.n = n;
#define SE2P_SPREAD_BOXES(NQV)                                                                     \
    hipLaunchKernelGGL((k_spread_strip_boxes<NQV, true>), dim3((unsigned)sp.off[n]), dim3(256),    \
                       lds, st, sp);                                                               \
    hipLaunchKernelGGL((k_spread_combine_boxes<NQV>), dim3((unsigned)cp.off[n]), dim3(64), 0, st, cp);
        switch (runs[0].gp.nq) {
        case 2: SE2P_SPREAD_BOXES(2); break;
        case 3: SE2P_SPREAD_BOXES(3); break;
        default: return fail(ctx, SE2P_ERR_ARG, "unsupported grid count %d", runs[0].gp.nq);
        }
#undef SE2P_SPREAD_BOXES
        SE2P_HIP(hipGetLastError());
    }
    return 0;
}

// Tile interpolation of many boxes, kPack boxes per launch (bitwise as
// single-box launches); the other boxes run one by one.
static int interp_boxes(se2p_ctx *ctx, std::vector<KRun> &runs)
{
    hipStream_t st = ctx->stream;
    std::vector<int> pk24, pk0;
    for (size_t k = 0; k < runs.size(); ++k) {
        const KRun &r = runs[k];
        if (r.L.itile && r.nrhs == 1 && !r.jac) (r.L.P == 24 ? pk24 : pk0).push_back((int)k);
        else if (int rc = krun_interp(ctx, r)) return rc;
    }
    for (int pt : {24, 0}) {
        const std::vector<int> &pk = pt == 24 ? pk24 : pk0;
        for (size_t s0 = 0; s0 < pk.size(); s0 += kPack) {
            const int n = (int)std::min<size_t>(kPack, pk.size() - s0);
            InterpPack ip{};
            size_t lds = 0;
            for (int j = 0; j < n; ++j) {
                const KRun &r = runs[(size_t)pk[s0 + (size_t)j]];
                ip.a[j] = interp_tile_args(r);
                ip.off[j + 1] = ip.off[j] + (int)r.gp.max_interp_chunks;
                lds = std::max(lds, r.L.lds_itile);
            }
            ip.n = n;
            if (pt == 24)
                hipLaunchKernelGGL((k_interp_tile_boxes<24>), dim3((unsigned)ip.off[n]), dim3(256), lds,
                                   st, ip);
            else
                hipLaunchKernelGGL((k_interp_tile_boxes<0>), dim3((unsigned)ip.off[n]), dim3(256), lds,
                                   st, ip);
            SE2P_HIP(hipGetLastError());
        }
    }
    return 0;
}

// Fourier-space sums of many independent boxes.  Every box gets a workspace
// region of its own, so that all of them are resident at once:
//   1. when every box's targets are its sources, one sort, gather and work
//      map sequence over all boxes on the context stream (mb_sort);
//      otherwise, on the stream slots, each box's sorts, gathers and work
//      maps;
//   2. on the context stream: spreading, boxes packed into shared launches;
//   3. on the slots: each box's transforms and filter;
//   4. on the context stream: interpolation, packed as the spreading.
// Each box is computed exactly as a single-box call computes it.
int kspace_sum_boxes(se2p_ctx *ctx, int kind, const se2p_box_desc *boxes, int nbox)
{
    std::vector<KRun> runs;
    std::vector<double> cost;
    for (int i = 0; i < nbox; ++i) {
        const se2p_box_desc &b = boxes[i];
        if (b.nt == 0) continue;
        if (b.ns == 0) {
            launch_zero(ctx->stream, b.u, 2 * b.nt);   // no sources: zero result
            continue;
        }
        const KGeom g{b.psrc, b.ptar, b.ns, b.nt, b.xi, b.eta, b.Lx, b.Ly, b.w, b.Mx, b.My, b.P};
        runs.emplace_back();
        krun_init(runs.back(), kind, g, b.f, b.n, 1, 2 * b.ns, 2 * b.nt, b.u, ctx->spread_chunk);
        cost.push_back((double)b.ns + (double)b.nt + 0.25 * (double)b.Mx * (double)b.My);
    }
    SE2P_HIP(hipGetLastError());
    const int nrun = (int)runs.size();
    if (nrun == 0) return 0;
    // one sort over all boxes when every box's targets are its sources
    MbSort ms;
    bool shared = nrun > 1;
    {
        int64_t n = 0, nk = 0;
        for (const KRun &r : runs) {
            shared = shared && r.same_points && r.nrhs == 1;
            n += r.g.ns;
            nk += r.L.nbins;
        }
        shared = shared && n <= 0x7fffffffLL && nk < 0x7fffffffLL;
        ms.n = (int)n;
        ms.nkeys = (int)nk;
    }
    for (KRun &r : runs) r.shared_sort = shared;
    std::vector<size_t> base((size_t)nrun);
    size_t total = 0;
    for (int k = 0; k < nrun; ++k) {
        Carve sizing;
        runs[(size_t)k].layout(sizing);
        base[(size_t)k] = total;
        total += sizing.off;
    }
    size_t mbase = total;
    if (shared) {
        Carve sizing;
        ms.layout(sizing, nrun, runs[0].ncomp);
        total += sizing.off;
    }
    if (int rc = ensure_workspace(ctx, total)) return rc;
    for (int k = 0; k < nrun; ++k) {
        Carve cv;
        cv.base = ctx->ws + base[(size_t)k];
        runs[(size_t)k].layout(cv);
    }
    if (shared) {
        Carve cv;
        cv.base = ctx->ws + mbase;
        ms.layout(cv, nrun, runs[0].ncomp);
        if (int rc = mb_sort(ctx, kind, runs, ms)) return rc;
    }
    // largest first, each to the slot with the least work so far
    const int nsl = std::min(kKspaceSlots, nrun);
    if (int rc = ensure_slots(ctx, nsl)) return rc;
    std::vector<int> order((size_t)nrun), slot_of((size_t)nrun);
    for (int k = 0; k < nrun; ++k) order[(size_t)k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return cost[(size_t)x] > cost[(size_t)y]; });
    {
        std::vector<double> load((size_t)nsl, 0.0);
        for (int k : order) {
            const int sl = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[(size_t)sl] += cost[(size_t)k];
            slot_of[(size_t)k] = sl;
        }
    }
    // per-box work on the slots, forked from and joined into the context stream
    auto on_slots = [&](auto &&work) {
        SE2P_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
        for (int k = 0; k < nsl; ++k)
            SE2P_HIP(hipStreamWaitEvent(ctx->slots[(size_t)k].st, ctx->fork_ev, 0));
        int rc = 0;
        for (int k : order) {
            SlotScope scope(ctx, slot_of[(size_t)k]);
            if ((rc = work(runs[(size_t)k])) != 0) break;
        }
        // join (also after an error, so that the context's stream stays ordered)
        for (int k = 0; k < nsl; ++k) {
            SE2P_HIP(hipEventRecord(ctx->slots[(size_t)k].done, ctx->slots[(size_t)k].st));
            SE2P_HIP(hipStreamWaitEvent(ctx->stream, ctx->slots[(size_t)k].done, 0));
        }
        return rc;
    };
    int rc = on_slots([&](KRun &r) {
        if (int e = get_plans(ctx, r.L.Mx, r.L.My, r.ncomp * r.nrp, 2 * r.nrp, &r.plans)) return e;
        if (shared) return 0;   // sorted, gathered and mapped above
        if (int e = krun_sort(ctx, r)) return e;
        krun_gather(ctx, r);
        SE2P_HIP(hipGetLastError());
        return 0;
    });
    if (rc) return rc;
    if ((rc = spread_boxes(ctx, runs)) != 0) return rc;
    rc = on_slots([&](KRun &r) {
        return transform_filter(ctx, r.L, r.g, r.plans, kind == 0 ? FILTER_SLP : FILTER_DLP, r.nrp,
                                r.grids, r.spec);
    });
    if (rc) return rc;
    return interp_boxes(ctx, runs);
}

// ---- adjoint -------------------------------------------------------------------
// Two passes, each one spreading + transform/filter + interpolation:
//   target pass: the forward grids (sources spread, filtered), interpolated at
//     the targets with derivative weights, weighted by du: dL/dptar;
//   source pass: du spread at the targets, transposed filter (the Stokeslet
//     filter is its own transpose), interpolated at the sources with value and
//     derivative weights: dL/df, dL/dn, dL/dpsrc.
template <int NR, int MODE>
static void launch_interp_adj(hipStream_t st, const KLayout &L, const GridPass &gp, bool tile,
                              const InterpAdjArgs &ia)
{
    if (tile && L.P == 24)
        hipLaunchKernelGGL((k_interp_adj_tile<NR, MODE, 24>), dim3((unsigned)gp.max_interp_chunks),
                           dim3(256), interp_adj_lds(L.P, NR), st, ia);
    else if (tile)
        hipLaunchKernelGGL((k_interp_adj_tile<NR, MODE, 0>), dim3((unsigned)gp.max_interp_chunks),
                           dim3(256), interp_adj_lds(L.P, NR), st, ia);
    else
        hipLaunchKernelGGL((k_interp_adj_direct<NR, MODE>), dim3((unsigned)((ia.n + 255) / 256)),
                           dim3(256), 0, st, ia);
}

// Targets are the sources, both sides wanted and no kept grids: the two
// passes share their points, so they run as one.  The densities' ncomp
// gridding columns and du's two are spread together (one sort and gather, one
// spreading of ncomp + 2 grids), one batched transform pair carries both
// filters (k_filter<FILTER_BOTH_*>), and the interpolation forms the target part from
// grids 0, 1 and the source part from grids 2.. .
static int kspace_adjoint_same(se2p_ctx *ctx, int kind, const KGeom &g, const double *f,
                               const double *nv, const double *du, double *dpsrc, double *dptar,
                               double *df, double *dn)
{
    hipStream_t st = ctx->stream;
    const int64_t n = g.ns;
    const int ncomp = kind == 0 ? 2 : 3, nq = ncomp + 2;
    KRun r;
    krun_init(r, kind, g, f, nv, 1, 2 * n, 2 * n, nullptr, ctx->spread_chunk);
    plan_pass(r.L, nq, 1, n, n, ctx->spread_chunk, &r.gp);
    // tile interpolation (and its work maps) where the staged grids fit
    const bool tile = interp_adj_lds(r.L.P, 2) <= kLdsMax;
    r.maps = r.gp.strip || tile;
    if (int rc = get_plans(ctx, r.L.Mx, r.L.My, nq, nq, &r.plans)) return rc;
    // the single-density run's buffers, with nq columns of gridding data,
    // grids and spectra
    r.ncomp = nq;
    Carve sizing;
    r.layout(sizing);
    if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
    Carve cv;
    cv.base = ctx->ws;
    r.layout(cv);
    r.ncomp = ncomp;

    stage_begin(ctx, ST_SORT);
    if (int rc = krun_sort(ctx, r)) return rc;
    stage_end(ctx, ST_SORT);

    stage_begin(ctx, ST_SPREAD);
    krun_gather(ctx, r);
    hipLaunchKernelGGL(k_gather_du_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, du,
                       r.kss.perm, (int)n, r.q + (int64_t)ncomp * n);
    if (int rc = spread(ctx, r.L, r.gp, r.sa)) return rc;
    stage_end(ctx, ST_SPREAD);

    if (int rc = transform_filter(ctx, r.L, g, r.plans, kind == 0 ? FILTER_BOTH_SLP : FILTER_BOTH_DLP,
                                  1, r.grids, r.spec))
        return rc;

    stage_begin(ctx, ST_INTERP);
    const KLayout &L = r.L;
    InterpAdjArgs a{};
    a.pts = g.psrc;
    a.perm = r.kss.perm;
    a.boff = r.kss.off;
    a.nbx = L.nbx;
    a.nby = L.nby;
    a.chunk_off = r.gp.ioffc;
    a.block_tile = r.gp.imap;
    a.ntx = L.ntx;
    a.nty = L.nty;
    a.n = n;
    a.grids = r.grids;
    a.gstride = L.G;
    a.Mx = L.Mx;
    a.My = L.My;
    a.P = L.P;
    a.Lx = g.Lx;
    a.Ly = g.Ly;
    a.h = L.h;
    a.c = L.c;
    a.w = g.w;
    a.ch2 = L.ch2;
    a.scale = L.interp_scale;
    a.du = du;
    a.f = f;
    a.nv = nv;
    // target part from grids 0, 1, source part from grids 2.. (one launch of
    // all ncomp + 2 grids holds too many registers per thread)
    InterpAdjArgs at = a, as = a;
    at.dpos = dptar;
    as.grids = r.grids + 2 * L.G;
    as.dpos = dpsrc;
    as.df = df;
    as.dn = dn;
    launch_interp_adj<2, ADJ_TARGET>(st, L, r.gp, tile, at);
    if (kind == 0)
        launch_interp_adj<2, ADJ_SOURCE_SLP>(st, L, r.gp, tile, as);
    else
        launch_interp_adj<3, ADJ_SOURCE_DLP>(st, L, r.gp, tile && interp_adj_lds(L.P, 3) <= kLdsMax, as);
    SE2P_HIP(hipGetLastError());
    stage_end(ctx, ST_INTERP);
    return 0;
}

int kspace_adjoint(se2p_ctx *ctx, int kind, const KGeom &g, const double *f, const double *nv,
                   const double *du, double *dpsrc, double *dptar, double *df, double *dn)
{
    hipStream_t st = ctx->stream;
    const int64_t ns = g.ns, nt = g.nt;
    const bool want_src = dpsrc || df || dn;
    if (ns == 0 || nt == 0) {
        // an empty side makes every grid zero
        launch_zero(st, dptar, 2 * nt);
        launch_zero(st, dpsrc, 2 * ns);
        launch_zero(st, df, 2 * ns);
        launch_zero(st, dn, 2 * ns);
        SE2P_HIP(hipGetLastError());
        return 0;
    }
    const KLayout L = make_layout(g);
    const int ncomp = kind == 0 ? 2 : 3;
    const bool same_points = (g.ptar == g.psrc) && (nt == ns);
    // the forward call's state, when it kept a matching one: bins, plans and
    // the filtered grids the target pass interpolates
    const KeptK *kept = same_points && ctx->kept_k && ctx->kept_k->match(kind, g, f, nv)
                            ? ctx->kept_k : nullptr;
    if (same_points && !kept && dptar && want_src && !ctx->split_adjoint)
        return kspace_adjoint_same(ctx, kind, g, f, nv, du, dpsrc, dptar, df, dn);
    GridPass gt, gs;   // target pass, source pass
    plan_pass(L, ncomp, 1, ns, nt, ctx->spread_chunk, &gt);
    plan_pass(L, 2, 1, nt, ns, ctx->spread_chunk, &gs);
    // the source pass spreads over the same bins with the same chunking as
    // the forward: its work plans are the forward's
    const bool share_maps = kept && gs.strip == kept->gp.strip && gs.chunk == kept->gp.chunk &&
                            gs.max_spread_chunks == kept->gp.max_spread_chunks &&
                            gs.max_interp_chunks == kept->gp.max_interp_chunks;
    if (kept) gt = kept->gp;
    const bool tile_t = interp_adj_lds(L.P, 2) <= kLdsMax;
    const bool tile_s = interp_adj_lds(L.P, ncomp) <= kLdsMax;
    FftPlans *pt = nullptr, *ps = nullptr;
    if (dptar && !kept)
        if (int rc = get_plans(ctx, L.Mx, L.My, ncomp, 2, &pt)) return rc;
    if (want_src)
        if (int rc = get_plans(ctx, L.Mx, L.My, 2, ncomp, &ps)) return rc;

    const int64_t nmax = std::max(ns, nt);
    KeySort kss, kst;
    int2 *anchor;
    double *fgg, *q, *grids;
    double2 *spec;
    auto layout = [&](Carve &cv) {
        if (!kept) {
            char *p = cv.take<char>(keysort_bytes(ns, L.nbins));
            if (p) keysort_carve(p, ns, L.nbins, &kss);
            if (!same_points) {
                p = cv.take<char>(keysort_bytes(nt, L.nbins));
                if (p) keysort_carve(p, nt, L.nbins, &kst);
            }
        }
        anchor = cv.take<int2>(nmax);
        fgg = cv.take<double>(3 * (size_t)nmax);
        q = cv.take<double>((size_t)ncomp * nmax);
        grids = cv.take<double>((size_t)ncomp * L.G);
        spec = cv.take<double2>((size_t)ncomp * L.GH);
        if (!kept) carve_pass(cv, L, &gt);
        if (share_maps) {
            GridPass keep_gs = gs;
            keep_gs.scnt = kept->gp.scnt;
            keep_gs.soffc = kept->gp.soffc;
            keep_gs.icnt = kept->gp.icnt;
            keep_gs.ioffc = kept->gp.ioffc;
            keep_gs.rng_start = kept->gp.rng_start;
            keep_gs.rng_len = kept->gp.rng_len;
            keep_gs.smap = kept->gp.smap;
            keep_gs.imap = kept->gp.imap;
            keep_gs.partial = gs.strip ? cv.take<double>((size_t)256 * gs.nq * gs.max_spread_chunks)
                                       : nullptr;
            gs = keep_gs;
        } else {
            carve_pass(cv, L, &gs);
        }
    };
    Carve sizing;
    layout(sizing);
    if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
    Carve cv;
    cv.base = ctx->ws;
    layout(cv);
    if (kept) kss = kept->kss;

    stage_begin(ctx, ST_SORT);
    if (!kept) {
        if (int rc = bin_points(ctx, L, g, g.psrc, ns, &kss)) return rc;
        if (!same_points)
            if (int rc = bin_points(ctx, L, g, g.ptar, nt, &kst)) return rc;
    }
    const KeySort &tbins = same_points ? kss : kst;
    SpreadArgs sat = spread_args(L, gt, kss.off, grids);     // sources -> grids
    SpreadArgs sas = spread_args(L, gs, tbins.off, grids);   // du at the targets -> grids
    if (dptar && !kept && (gt.strip || tile_t)) tile_maps(st, L, gt, sat, tbins.off);
    if (want_src && !share_maps && (gs.strip || tile_s)) tile_maps(st, L, gs, sas, kss.off);
    stage_end(ctx, ST_SORT);

    InterpAdjArgs ia{};
    ia.nbx = L.nbx;
    ia.nby = L.nby;
    ia.ntx = L.ntx;
    ia.nty = L.nty;
    ia.grids = grids;
    ia.gstride = L.G;
    ia.Mx = L.Mx;
    ia.My = L.My;
    ia.P = L.P;
    ia.Lx = g.Lx;
    ia.Ly = g.Ly;
    ia.h = L.h;
    ia.c = L.c;
    ia.w = g.w;
    ia.ch2 = L.ch2;
    ia.scale = L.interp_scale;

    if (dptar && kept && kept->jac) {
        // the forward's output gradients at the targets, contracted with du
        stage_begin(ctx, ST_INTERP);
        hipLaunchKernelGGL(k_target_grad, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st,
                           kept->jac, du, nt, dptar);
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_INTERP);
    } else if (dptar && kept) {
        // the forward's filtered grids, interpolated with derivative weights
        stage_begin(ctx, ST_INTERP);
        InterpAdjArgs a = ia;
        a.grids = kept->grids;
        a.pts = g.ptar;
        a.perm = tbins.perm;
        a.boff = tbins.off;
        a.chunk_off = gt.ioffc;
        a.block_tile = gt.imap;
        a.n = nt;
        a.du = du;
        a.dpos = dptar;
        launch_interp_adj<2, ADJ_TARGET>(st, L, gt, tile_t, a);
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_INTERP);
    } else if (dptar) {
        stage_begin(ctx, ST_SPREAD);
        const unsigned blocks = (unsigned)((ns + 255) / 256);
        if (kind == 0)
            hipLaunchKernelGGL(k_gather_grid_src<0>, dim3(blocks), dim3(256), 0, st, g.psrc, f, nv,
                               kss.perm, (int)ns, 1, 1, (int64_t)0, g.Lx, g.Ly, L.h, L.P, L.Mx,
                               L.My, L.c, g.w, anchor, fgg, q);
        else
            hipLaunchKernelGGL(k_gather_grid_src<1>, dim3(blocks), dim3(256), 0, st, g.psrc, f, nv,
                               kss.perm, (int)ns, 1, 1, (int64_t)0, g.Lx, g.Ly, L.h, L.P, L.Mx,
                               L.My, L.c, g.w, anchor, fgg, q);
        sat.s.anchor = anchor;
        sat.s.fgg = fgg;
        sat.s.q = q;
        sat.s.ns = ns;
        if (int rc = spread(ctx, L, gt, sat)) return rc;
        stage_end(ctx, ST_SPREAD);
        if (int rc = transform_filter(ctx, L, g, pt, kind == 0 ? FILTER_SLP : FILTER_DLP, 1, grids, spec))
            return rc;
        stage_begin(ctx, ST_INTERP);
        InterpAdjArgs a = ia;
        a.pts = g.ptar;
        a.perm = tbins.perm;
        a.boff = tbins.off;
        a.chunk_off = gt.ioffc;
        a.block_tile = gt.imap;
        a.n = nt;
        a.du = du;
        a.dpos = dptar;
        launch_interp_adj<2, ADJ_TARGET>(st, L, gt, tile_t, a);
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_INTERP);
    }
    if (want_src) {
        stage_begin(ctx, ST_SPREAD);
        hipLaunchKernelGGL(k_gather_grid_src<0>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st,
                           g.ptar, du, nullptr, tbins.perm, (int)nt, 1, 1, (int64_t)0, g.Lx, g.Ly,
                           L.h, L.P, L.Mx, L.My, L.c, g.w, anchor, fgg, q);
        sas.s.anchor = anchor;
        sas.s.fgg = fgg;
        sas.s.q = q;
        sas.s.ns = nt;
        if (int rc = spread(ctx, L, gs, sas)) return rc;
        stage_end(ctx, ST_SPREAD);
        if (int rc = transform_filter(ctx, L, g, ps, kind == 0 ? FILTER_SLP : FILTER_DLP_ADJ, 1, grids, spec))
            return rc;
        stage_begin(ctx, ST_INTERP);
        InterpAdjArgs a = ia;
        a.pts = g.psrc;
        a.perm = kss.perm;
        a.boff = kss.off;
        a.chunk_off = gs.ioffc;
        a.block_tile = gs.imap;
        a.n = ns;
        a.f = f;
        a.nv = nv;
        a.dpos = dpsrc;
        a.df = df;
        a.dn = dn;
        if (kind == 0) launch_interp_adj<2, ADJ_SOURCE_SLP>(st, L, gs, tile_s, a);
        else launch_interp_adj<3, ADJ_SOURCE_DLP>(st, L, gs, tile_s, a);
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_INTERP);
    }
    return 0;
}

}  // namespace se2p
