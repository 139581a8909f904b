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
__global__ void k_bin_keys(const double *__restrict__ p, int n, double Lx, double Ly, double h,
                           int P, int Mx, int My, int nby, int *__restrict__ keys,
                           int *__restrict__ slot, int *__restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int mx, my;
    double px, py;
    se2p_anchor(p[2 * i], p[2 * i + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
    int ax = ((mx % Mx) + Mx) % Mx, ay = ((my % My) + My) % My;
    const int k = (ax / kBin) * nby + ay / kBin;
    keys[i] = k;
    slot[i] = atomicAdd(&count[k], 1);
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

// kind 0: q = f (2 comps); kind 1: q = (n1 f1, n2 f1 + n1 f2, n2 f2)
template <int KIND>
__global__ void k_gather_grid_src(const double *__restrict__ p, const double *__restrict__ f,
                                  const double *__restrict__ nv, const int *__restrict__ perm,
                                  int n, int r0, int nr, int64_t dstride, double Lx, double Ly,
                                  double h, int P, int Mx, int My, double c, double w,
                                  int2 *__restrict__ anchor, double *__restrict__ fgg,
                                  double *__restrict__ q)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    int mx, my;
    double px, py;
    se2p_anchor(p[2 * o], p[2 * o + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
    anchor[i] = make_int2(((mx % Mx) + Mx) % Mx, ((my % My) + My) % My);
    double bx, e3x, e3y;
    fgg_factors(px, py, c, w, h, &bx, &e3x, &e3y);
    fgg[i] = bx;
    fgg[n + i] = e3x;
    fgg[2 * n + i] = e3y;
    if (KIND == 0) {
        for (int r = 0; r < nr; ++r) {
            const double *fr = f + (int64_t)(r0 + r) * dstride;
            q[(int64_t)(2 * r) * n + i] = fr[2 * o];
            q[(int64_t)(2 * r + 1) * n + i] = fr[2 * o + 1];
        }
    } else {
        double n1 = nv[2 * o], n2 = nv[2 * o + 1];
        for (int r = 0; r < nr; ++r) {
            const double *fr = f + (int64_t)(r0 + r) * dstride;
            double f1 = fr[2 * o], f2 = fr[2 * o + 1];
            q[(int64_t)(3 * r) * n + i] = n1 * f1;
            q[(int64_t)(3 * r + 1) * n + i] = n2 * f1 + n1 * f2;
            q[(int64_t)(3 * r + 2) * n + i] = n2 * f2;
        }
    }
}

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
};

// Cyclic list of anchor bins whose anchors a satisfy (t - a) mod M in [0, P]
// for some node t of [t0, t0 + kTile): anchors in [t0 - P, t0 + kTile - 1].
__device__ inline int bins_for_tile(int t0, int M, int P, int nb, int *out)
{
    int n = 0;
    int span = kTile + P;
    if (span > M) span = M;
    int start = t0 - P;
    start = ((start % M) + M) % M;
    for (int d = 0; d < span; ++d) {
        int a = start + d;
        if (a >= M) a -= M;
        int b = a / kBin;
        bool seen = false;
        for (int k = 0; k < n; ++k) seen |= (out[k] == b);
        if (!seen) out[n++] = b;
    }
    (void)nb;
    return n;
}

template <int NQ, bool WRAP>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs a)
{
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
        int bxl[32], byl[32];
        int nbxl = bins_for_tile(X0, a.Mx, a.P, a.nbx, bxl);
        int nbyl = bins_for_tile(Y0, a.My, a.P, a.nby, byl);
        int nr = 0, tot = 0;
        for (int i = 0; i < nbxl; ++i) {
            // y bins for this x bin form runs of consecutive keys
            int k = 0;
            while (k < nbyl) {
                int b0 = byl[k], len = 1;
                while (k + len < nbyl && byl[k + len] == b0 + len) ++len;
                int key0 = bxl[i] * a.nby + b0;
                int s0 = a.boff[key0], s1 = a.boff[key0 + len];
                if (s1 > s0 && nr < kMaxRanges) {
                    rstart[nr] = s0;
                    rpref[nr] = tot;
                    tot += s1 - s0;
                    ++nr;
                }
                k += len;
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
                    for (int c = 0; c < NQ; ++c) Q[sl * NQ + c] = a.s.q[(int64_t)c * a.s.ns + k];
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
        for (int c = 0; c < NQ; ++c) a.grids[c * a.gstride + idx] = acc[c];
    }
}

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

// SLP: input grids [2r], [2r+1] -> output in place.
__global__ void k_filter_slp(FilterArgs a)
{
    const int Mh = a.My / 2 + 1;
    const int64_t n = (int64_t)a.Mx * Mh;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int j = (int)(p / Mh), k = (int)(p % Mh);
    const double k1 = wavenumber(j, a.Mx, a.Lx);
    const double k2 = 2.0 * SE2P_PI / a.Ly * k;
    const double K2 = k1 * k1 + k2 * k2;
    double e = 0.0, a11 = 0.0, a22 = 0.0, a12 = 0.0;
    if (p != 0) {
        // mex_stokes_slp_kspace.cpp:200
        e = (1.0 / (K2 * K2) + 0.25 / (K2 * a.xi * a.xi)) *
            exp(-0.25 * (1 - a.eta) / (a.xi * a.xi) * K2) * a.scale;
        const bool nyx = (2 * j == a.Mx), nyy = (2 * k == a.My);
        a11 = k2 * k2;
        a22 = k1 * k1;
        a12 = (nyx != nyy) ? 0.0 : k1 * k2;   // Hermitian part on Nyquist lines
    }
    for (int r = 0; r < a.nrhs; ++r) {
        double2 *H1 = a.spec + (int64_t)(2 * r) * a.sstride;
        double2 *H2 = H1 + a.sstride;
        const double2 h1 = H1[p], h2 = H2[p];
        H1[p] = make_double2((a11 * h1.x - a12 * h2.x) * e, (a11 * h1.y - a12 * h2.y) * e);
        H2[p] = make_double2((a22 * h2.x - a12 * h1.x) * e, (a22 * h2.y - a12 * h1.y) * e);
    }
}

// DLP: input grids [3r], [3r+1], [3r+2] (A = n1f1, B = n2f1+n1f2, C = n2f2)
// -> output grids [2r], [2r+1], multiplied by -i so that the real inverse
// transform yields the imaginary part of the reference's inverse.
__global__ void k_filter_dlp(FilterArgs a)
{
    const int Mh = a.My / 2 + 1;
    const int64_t n = (int64_t)a.Mx * Mh;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int j = (int)(p / Mh), k = (int)(p % Mh);
    const double k1 = wavenumber(j, a.Mx, a.Lx);
    const double k2 = 2.0 * SE2P_PI / a.Ly * k;
    const double K2 = k1 * k1 + k2 * k2;
    double et = 0.0, b111 = 0.0, b112 = 0.0, b122 = 0.0, b222 = 0.0;
    if (p != 0) {
        // mex_stokes_dlp_kspace.cpp:232-247
        const double e = exp(-0.25 * (1 - a.eta) / (a.xi * a.xi) * K2) / K2 / K2;
        const double t = 1 + 0.25 * K2 / a.xi / a.xi;
        et = e * t * a.scale;
        const bool nyx = (2 * j == a.Mx), nyy = (2 * k == a.My);
        // anti-Hermitian part: terms odd in k1 vanish on the x Nyquist line,
        // terms odd in k2 on the y Nyquist line
        if (!nyx) {
            b111 = 3.0 * k1 * K2 - 2.0 * k1 * k1 * k1;
            b122 = k1 * K2 - 2.0 * k1 * k2 * k2;
        }
        if (!nyy) {
            b112 = k2 * K2 - 2.0 * k1 * k1 * k2;
            b222 = 3.0 * k2 * K2 - 2.0 * k2 * k2 * k2;
        }
    }
    for (int r = 0; r < a.nrhs; ++r) {
        const double2 A = a.spec[(int64_t)(3 * r) * a.sstride + p];
        const double2 B = a.spec[(int64_t)(3 * r + 1) * a.sstride + p];
        const double2 C = a.spec[(int64_t)(3 * r + 2) * a.sstride + p];
        const double o1r = et * (b111 * A.x + b112 * B.x + b122 * C.x);
        const double o1i = et * (b111 * A.y + b112 * B.y + b122 * C.y);
        const double o2r = et * (b112 * A.x + b122 * B.x + b222 * C.x);
        const double o2i = et * (b112 * A.y + b122 * B.y + b222 * C.y);
        // all threads read their own element before writing it; grid 2r and
        // 2r+1 never alias an unread input of another thread
        a.spec[(int64_t)(2 * r) * a.sstride + p] = make_double2(o1i, -o1r);
        a.spec[(int64_t)(2 * r + 1) * a.sstride + p] = make_double2(o2i, -o2r);
    }
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

__global__ void k_zero(double *__restrict__ u, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = 0.0;
}

// ---- rocFFT plans ------------------------------------------------------------
static int get_plans(se2p_ctx *ctx, int Mx, int My, int nfwd, int ninv, FftPlans **out)
{
    auto key = std::make_tuple(Mx, My, nfwd, ninv);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = &it->second;
        return 0;
    }
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

// ---- host driver -------------------------------------------------------------
int kspace_sum(se2p_ctx *ctx, int kind, const KGeom &g, const double *f, const double *nv,
               int nrhs, int64_t dstride, int64_t ustride, double *u)
{
    hipStream_t st = ctx->stream;
    if (g.nt == 0 || nrhs == 0) return 0;
    const int64_t ns = g.ns, nt = g.nt;
    const int Mx = g.Mx, My = g.My, P = g.P;
    const double h = g.Lx / Mx;                  // also used for y (:45)
    const double c = -2 * g.xi * g.xi / g.eta;
    const int ncomp = kind == 0 ? 2 : 3;
    // right-hand sides per pass (grids per pass = ncomp * RB)
    const int RB = (nrhs >= 4 && nrhs % 4 == 0) ? 4 : 1;
    const int nq = ncomp * RB;
    const int64_t G = (int64_t)Mx * My;
    const int Mh = My / 2 + 1;
    const int64_t GH = (int64_t)Mx * Mh;
    const int nbx = (Mx + kBin - 1) / kBin, nby = (My + kBin - 1) / kBin;
    const int nbins = nbx * nby;

    FftPlans *plans = nullptr;
    if (ns > 0)
        if (int rc = get_plans(ctx, Mx, My, nq, 2 * RB, &plans)) return rc;

    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t ks = keysort_bytes(ns, nbins);
    const size_t nsz = (size_t)(ns > 0 ? ns : 1);
    size_t need = ks + up(sizeof(int2) * nsz) + up(sizeof(double) * 3 * nsz) +
                  up(sizeof(double) * nq * nsz) + up(sizeof(double) * nq * (size_t)G) +
                  up(sizeof(double2) * nq * (size_t)GH);
    if (int rc = ensure_workspace(ctx, need)) return rc;
    char *base = ctx->ws;
    size_t o = 0;
    KeySort kss;
    keysort_carve(base + o, ns, nbins, &kss); o += ks;
    int2 *anchor = (int2 *)(base + o); o += up(sizeof(int2) * nsz);
    double *fgg = (double *)(base + o); o += up(sizeof(double) * 3 * nsz);
    double *q = (double *)(base + o); o += up(sizeof(double) * nq * nsz);
    double *grids = (double *)(base + o); o += up(sizeof(double) * nq * (size_t)G);
    double2 *spec = (double2 *)(base + o); o += up(sizeof(double2) * nq * (size_t)GH);

    const int threads = 256;
    if (ns == 0) {
        // no sources: every grid is zero, so is the result
        for (int r = 0; r < nrhs; ++r)
            hipLaunchKernelGGL(k_zero, dim3((unsigned)((2 * nt + threads - 1) / threads)),
                               dim3(threads), 0, st, u + (int64_t)r * ustride, 2 * nt);
        SE2P_HIP(hipGetLastError());
        return 0;
    }

    stage_begin(ctx, ST_SORT);
    keysort_clear(st, &kss, nbins);
    hipLaunchKernelGGL(k_bin_keys, dim3((unsigned)((ns + threads - 1) / threads)), dim3(threads), 0,
                       st, g.psrc, (int)ns, g.Lx, g.Ly, h, P, Mx, My, nby, kss.key, kss.slot,
                       kss.count);
    if (int rc = keysort_finish(ctx, &kss, ns, nbins, true)) return rc;
    stage_end(ctx, ST_SORT);

    const bool wrap = (Mx <= P) || (My <= P);
    const size_t lds = sizeof(double) * ((size_t)2 * kBatch * (P + 1) + (size_t)kBatch * nq + (P + 1)) +
                       sizeof(int) * 2 * kBatch;
    const double ch2 = c * h * h;
    double s_int = 4 * g.xi * g.xi / g.eta;                  // :325-328
    s_int = s_int * s_int * h * h / SE2P_PI;
    const double interp_scale = s_int / (4 * SE2P_PI);

    for (int r0 = 0; r0 < nrhs; r0 += RB) {
        stage_begin(ctx, ST_SPREAD);
        if (kind == 0)
            hipLaunchKernelGGL(k_gather_grid_src<0>, dim3((unsigned)((ns + threads - 1) / threads)),
                               dim3(threads), 0, st, g.psrc, f, nv, kss.perm, (int)ns, r0, RB,
                               dstride, g.Lx, g.Ly, h, P, Mx, My, c, g.w, anchor, fgg, q);
        else
            hipLaunchKernelGGL(k_gather_grid_src<1>, dim3((unsigned)((ns + threads - 1) / threads)),
                               dim3(threads), 0, st, g.psrc, f, nv, kss.perm, (int)ns, r0, RB,
                               dstride, g.Lx, g.Ly, h, P, Mx, My, c, g.w, anchor, fgg, q);
        SpreadArgs sa;
        sa.s.anchor = anchor;
        sa.s.fgg = fgg;
        sa.s.q = q;
        sa.s.ns = ns;
        sa.boff = kss.off;
        sa.nbx = nbx;
        sa.nby = nby;
        sa.Mx = Mx;
        sa.My = My;
        sa.P = P;
        sa.ch2 = ch2;
        sa.grids = grids;
        sa.gstride = G;
        dim3 sg((Mx + kTile - 1) / kTile, (My + kTile - 1) / kTile);
#define SE2P_SPREAD(NQV)                                                                        \
    if (wrap) hipLaunchKernelGGL((k_spread<NQV, true>), sg, dim3(256), lds, st, sa);            \
    else hipLaunchKernelGGL((k_spread<NQV, false>), sg, dim3(256), lds, st, sa);
        switch (nq) {
        case 2: SE2P_SPREAD(2); break;
        case 3: SE2P_SPREAD(3); break;
        case 8: SE2P_SPREAD(8); break;
        case 12: SE2P_SPREAD(12); break;
        default: return fail(ctx, SE2P_ERR_ARG, "unsupported grid count %d", nq);
        }
#undef SE2P_SPREAD
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_SPREAD);

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
        fa.sstride = GH;
        fa.Mx = Mx;
        fa.My = My;
        fa.nrhs = RB;
        fa.Lx = g.Lx;
        fa.Ly = g.Ly;
        fa.xi = g.xi;
        fa.eta = g.eta;
        fa.scale = 1.0 / ((double)Mx * (double)My);
        const unsigned fblocks = (unsigned)((GH + threads - 1) / threads);
        if (kind == 0) hipLaunchKernelGGL(k_filter_slp, dim3(fblocks), dim3(threads), 0, st, fa);
        else hipLaunchKernelGGL(k_filter_dlp, dim3(fblocks), dim3(threads), 0, st, fa);
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

        stage_begin(ctx, ST_INTERP);
        InterpArgs ia;
        ia.ptar = g.ptar;
        ia.nt = nt;
        ia.grids = grids;
        ia.gstride = G;
        ia.Mx = Mx;
        ia.My = My;
        ia.P = P;
        ia.nr = RB;
        ia.Lx = g.Lx;
        ia.Ly = g.Ly;
        ia.h = h;
        ia.c = c;
        ia.w = g.w;
        ia.ch2 = ch2;
        ia.scale = interp_scale;
        ia.u = u + (int64_t)r0 * ustride;
        ia.ustride = ustride;
        const unsigned iblocks = (unsigned)((nt + threads - 1) / threads);
        if (RB == 1) hipLaunchKernelGGL(k_interp<1>, dim3(iblocks), dim3(threads), 0, st, ia);
        else hipLaunchKernelGGL(k_interp<4>, dim3(iblocks), dim3(threads), 0, st, ia);
        SE2P_HIP(hipGetLastError());
        stage_end(ctx, ST_INTERP);
    }
    return 0;
}

}  // namespace se2p
