// se2p_real.hip -- real-space part of the spectral Ewald sum (Stokeslet SLP
// and stresslet DLP) on cell lists.
//
// Behaviour follows mex_stokes_slp_real.cpp:58-346 and
// mex_stokes_dlp_real.cpp:15-211:
//   * the box is split into nside_x x nside_y cells; every target sums over
//     ALL sources of its own cell (no cutoff) and over the sources of the 8
//     neighbour cells (periodic image offset) with r^2 < Lx*Ly/(nsx*nsy);
//   * SLP adds -(gamma/2 + 1 + ln xi) f for coincident points (r^2 == 0), DLP
//     skips them;
//   * the result is divided by 4 pi and written in the caller's target order.
//
// GPU mapping.  Sources are sorted by cell, stable (a cell keeps input
// order).  Targets are sorted by cell and, inside a cell, by a 4x4 Morton
// sub-cell, so the 64 targets of a chunk sit close together: a neighbour-cell
// source then tends to be inside the cutoff for all of them or for none, and
// the wavefront skips the pair code when none passes.  A 256-thread
// workgroup takes one chunk of up to 64 targets of one cell; its four
// wavefronts hold the same 64 targets (one per lane) and split every source
// cell four ways (wave w takes sources w, w+4, ...).  Source data is
// wave-uniform and arrives through the scalar cache.  The four partial sums
// are added in a fixed order through LDS, so results are deterministic (no
// atomics).  E1 for the Stokeslet comes from a piecewise polynomial table in
// LDS, exp(-x) from a 256-entry table (se2p_exp_neg).
//
// Neighbour cells are pruned by sub-cell: the bounding box of every
// (cell, sub-cell) run of the sorted sources is computed once per call, and a
// workgroup skips every run whose box (at the image offset) is farther than
// the cutoff from the bounding box of its 64 targets.  Skipped pairs all fail
// the cutoff test, so the sum is unchanged; the test has a relative margin of
// 1e-9 on the squared distance, far above the rounding of either side.
//
// Workgroups are numbered so that each of the 8 XCDs (which receive
// workgroups round-robin) takes a contiguous band of cells: its L2 then holds
// only the sources of its band and a one-cell halo.
#include <cstring>
#include <mutex>
#include <vector>

#include "se2p_device.h"
#include "se2p_fsort.h"
#include "se2p_internal.h"

namespace se2p {

static __constant__ double c_e1_table[SE2P_E1_NPIECE * SE2P_E1_ROW];
// 2^(j/256), the table of se2p_exp_neg; each workgroup copies it to LDS
static __constant__ double c_exp2_table[SE2P_EXP2_N] = SE2P_EXP2_TABLE;

// ---- fixed scratch of the no-allocation entry point (se2p_dlp_real_noalloc)
// The sorted state of that entry point (cell lists, source records, chunk
// plans, run boxes) is carved from this array instead of the context's
// workspace.  Each GPU holds its own copy from module load until the process
// ends.  On the host the name is only a stand-in for the symbol, not a device
// address: noalloc_scratch_init looks up each device's address once, with
// that device current, and host code uses only that address afterwards.
static __device__ __align__(256) char scratch[SE2P_NOALLOC_SCRATCH_BYTES];

constexpr int kScratchDevices = 64;

int noalloc_scratch_init(se2p_ctx *ctx)
{
    static std::mutex mu;
    static char *addr[kScratchDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    const int d = ctx->device;
    if (d < 0 || d >= kScratchDevices)
        return fail(ctx, SE2P_ERR_HIP, "no fixed scratch for device %d (devices 0..%d)", d,
                    kScratchDevices - 1);
    if (!addr[d]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();   // leave no sticky error for later calls
            return fail(ctx, SE2P_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        }
        addr[d] = (char *)p;
    }
    ctx->scratch = addr[d];
    return 0;
}

constexpr int kSubLog = 2;   // 4x4 sub-cells per cell for target ordering
constexpr int kWaves = 4;

// Neighbour j = 0..7 in the reference's order (mex_stokes_slp_real.cpp:61-62):
// the 3x3 block in x-major order with the centre left out.
__device__ __forceinline__ int nb_dx(int j) { return (j < 4 ? j : j + 1) / 3 - 1; }
__device__ __forceinline__ int nb_dy(int j) { return (j < 4 ? j : j + 1) % 3 - 1; }

// 1/x from the hardware estimate and one Newton step (|x| normal, finite).
__device__ __forceinline__ double rcp_nr(double x)
{
    const double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Per-box parameters of the real-space sum, read by the kernels from device
// memory: one record per independent periodic box (a single-box call has
// one).  Points of all boxes are numbered globally: box b owns sources
// [s_base, s_base + ns), targets [t_base, t_base + nt) and cells
// [cell_base, cell_base + nsx*nsy).
struct RealBox {
    const double *psrc, *ptar, *f, *nv;   // caller arrays (nv: stresslet)
    double *u;                            // result r at u + r*ustride, caller's order
    int ns, nt;
    int s_base, t_base, cell_base;
    int nsx, nsy;
    double Lx, Ly, xi, xi2, cut2, self;
};

// Up to kBoxesPerLaunch box records by value (a kernel argument of about
// 2 KB) and the kernel that writes them, with the box offset arrays
// [s_base..., ns | t_base..., nt | cell_base..., ncell], to the device.
constexpr int kBoxesPerLaunch = 16;
struct BoxBatch {
    RealBox b[kBoxesPerLaunch];
    int first, count;
};
static_assert(sizeof(BoxBatch) < 3072, "kernel argument size");

__global__ void k_write_boxes(BoxBatch bb, RealBox *__restrict__ tab, int *__restrict__ offs,
                              int nbox, int ns, int nt, int ncell)
{
    const int i = threadIdx.x;
    if (i < bb.count) {
        const int g = bb.first + i;
        tab[g] = bb.b[i];
        offs[g] = bb.b[i].s_base;
        offs[nbox + 1 + g] = bb.b[i].t_base;
        offs[2 * (nbox + 1) + g] = bb.b[i].cell_base;
    }
    if (i == 0 && bb.first + bb.count == nbox) {
        offs[nbox] = ns;
        offs[2 * nbox + 1] = nt;
        offs[3 * nbox + 2] = ncell;
    }
}

// The box records as kernels see them: a device table for several boxes, or
// the one record by value (no device copy, so a single-box call stays
// capturable and needs no extra launch).
struct BoxTable {
    const RealBox *tab;   // [nbox] on the device, or null: one box, `one`
    RealBox one;
    __device__ __forceinline__ const RealBox &get(int i) const { return tab ? tab[i] : one; }
};

struct RealArgs {
    const double2 *tpos;   // targets in (cell, sub-cell) order
    const int *tperm;      // that order -> global target index
    const int *toff;       // [ncell*16+1] run offsets of (cell, sub-cell) keys
    const int *chunk_off;  // [ncell+1] exclusive scan of ceil(targets/64)
    const int *block_cell; // chunk -> cell
    const double2 *spos;   // sources in cell order
    const double *q;       // densities in cell order, [nrhs][ncomp][ns]
    const int *soff;       // [ncell*16+1] run offsets of the sources' (cell, sub-cell) keys
    int64_t ns;            // all sources (density stride)
    int ncell;             // all cells
    int nrhs;
    int r_begin, r_end;    // right-hand sides handled by this launch
    int64_t ustride;
    BoxTable boxes;
    const int *cell_box;   // cell -> box (null: one box)
    unsigned long long *pair_counts;
};

// ---- pair kernels --------------------------------------------------------
// E1(x) as se2p_e1 computes it (same table, same operations), reading the
// coefficients from LDS in groups of four: a scheduling barrier after each
// group keeps the compiler from holding all fifteen in registers at once,
// which would cost the Stokeslet kernels a third of their occupancy.
__device__ __forceinline__ double e1_lds(double x, double emx, const double *tab)
{
    const double *row = tab + SE2P_E1_ROW * se2p_e1_piece(x);
    double s = x * row[0] + row[1];
    double v = row[2 + SE2P_E1_NCOEF - 1];
#pragma unroll
    for (int k = SE2P_E1_NCOEF - 2; k >= 0; --k) {
        v = v * s + row[2 + k];
        if (k % 4 == 0) __builtin_amdgcn_sched_barrier(0);
    }
    if (x < 1.0) return x * v - SE2P_EULER_GAMMA - log(x);
    return emx * v;
}

// Stokeslet pair, before 1/4pi: 0.5 E1(xi^2 r^2) f + e^{-xi^2 r^2}((d.f) d/r^2 - f),
// arranged as in mex_stokes_slp_real.cpp:207-225: c = d1 f2 - d2 f1,
// u1 += E/2 f1 + (e/r2) d2 c,  u2 += E/2 f2 - (e/r2) d1 c.
// Stokeslet pair geometry factors: hE = E1(xi^2 r^2)/2, g = e^{-xi^2 r^2}/r^2.
__device__ __forceinline__ void slp_factors(double r2, double xi, const double *e1tab,
                                            const double *etab, double &hE, double &g)
{
    const double x = r2 * (xi * xi);   // xi^2 is loop-invariant: one multiply per pair
    const double e = se2p_exp_neg(x, etab);
    hE = 0.5 * e1_lds(x, e, e1tab);
    g = e * rcp_nr(r2);
}

// one density (f1, f2): u1 += E/2 f1 + (e/r2) d2 c, u2 += E/2 f2 - (e/r2) d1 c
__device__ __forceinline__ void slp_add(double d1, double d2, double hE, double g, double f1,
                                        double f2, double &a1, double &a2)
{
    const double c = d1 * f2 - d2 * f1;
    a1 = fma(hE, f1, fma(g * d2, c, a1));
    a2 = fma(hE, f2, fma(-(g * d1), c, a2));
}

template <int RB>
__device__ __forceinline__ void slp_pair(double d1, double d2, double r2, double xi,
                                         const double *e1tab, const double *etab, const double *q,
                                         int64_t qs, double (&acc)[RB][2])
{
    double hE, g;
    slp_factors(r2, xi, e1tab, etab, hE, g);
#pragma unroll
    for (int r = 0; r < RB; ++r)
        slp_add(d1, d2, hE, g, q[(2 * r) * qs], q[(2 * r + 1) * qs], acc[r][0], acc[r][1]);
}

// One right-hand side: each source is one record of SrcRec<KIND>::n doubles,
// fetched by one scalar load: (x, y, f1, f2) for the Stokeslet; for the
// stresslet (x, y, -4 s0, -4 s1, -4 s2, b c1, b c2, b s1) with
// s = (S11, S12 + S21, S22), c1 = 3 s0 + s2, c2 = s0 + 3 s2, b = 2 xi^2.
template <int KIND>
struct SrcRec {
    static constexpr int n = KIND == 0 ? 4 : 8;
};

// q holds, per right-hand side, the five source values
// (S11, S12+S21, S22, 3 S11 + S22, S11 + 3 S22) (kDlpComp).
constexpr int kDlpComp = 5;

// Stresslet pair geometry factors: g = e2 fb, pe = 2 xi^2 e2.
__device__ __forceinline__ void dlp_factors(double r2, double xi2, const double *etab, double &g,
                                            double &pe)
{
    const double x = xi2 * r2;
    const double e2 = se2p_exp_neg(x, etab);
    const double ir2 = rcp_nr(r2);
    g = -4.0 * e2 * (1.0 + x) * (ir2 * ir2);
    pe = 2.0 * xi2 * e2;
}

// one density: u_i += g d_i (d.S d) + pe l_i
__device__ __forceinline__ void dlp_add(double d1, double d2, double d11, double d12, double d22,
                                        double g, double pe, double s0, double s1, double s2,
                                        double c1, double c2, double &a1, double &a2)
{
    const double gq = g * (d11 * s0 + d12 * s1 + d22 * s2);
    a1 += gq * d1 + pe * (d1 * c1 + d2 * s1);
    a2 += gq * d2 + pe * (d1 * s1 + d2 * c2);
}

template <int RB>
__device__ __forceinline__ void dlp_pair(double d1, double d2, double d11, double d22, double r2,
                                         double xi2, const double *etab, const double *q,
                                         int64_t qs, double (&acc)[RB][2])
{
    double g, pe;
    dlp_factors(r2, xi2, etab, g, pe);
    const double d12 = d1 * d2;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const double *qr = q + (kDlpComp * r) * qs;
        dlp_add(d1, d2, d11, d12, d22, g, pe, qr[0], qr[qs], qr[2 * qs], qr[3 * qs], qr[4 * qs],
                acc[r][0], acc[r][1]);
    }
}

// Difference vector with the reference's rounding order:
//   SLP: d = src - (tgt - z)   (mex_stokes_slp_real.cpp:155,256-260)
//   DLP: d = (src - tgt) + z   (mex_stokes_dlp_real.cpp:117-118,162-163)
// With z = 0 both reduce exactly to the own-cell forms (the SLP sign flip is
// exact and the pair term is even in d).
template <int KIND>
__device__ __forceinline__ void diff(double sx, double sy, double tx, double ty, double zx,
                                     double zy, double &d1, double &d2)
{
    if (KIND == 0) {
        d1 = sx - (tx - zx);
        d2 = sy - (ty - zy);
    } else {
        d1 = (sx - tx) + zx;
        d2 = (sy - ty) + zy;
    }
}

// ---- work order and pruning helpers ---------------------------------------
// Workgroup -> chunk such that XCD x (workgroups x, x+8, ...) takes chunks
// [x*per, (x+1)*per).  The launch grid is a multiple of 8 and at least the
// chunk count.  Returns -1 for workgroups without a chunk.
__device__ __forceinline__ int xcd_chunk(int nvalid)
{
    const int per = (nvalid + 7) >> 3;
    const int slot = blockIdx.x >> 3;
    const int b = (blockIdx.x & 7) * per + slot;
    return (slot < per && b < nvalid) ? b : -1;
}

// Bounding box of the wavefront's active lanes.
__device__ __forceinline__ void wave_bbox(bool active, double x, double y, double &x0, double &y0,
                                          double &x1, double &y1)
{
    const double inf = __builtin_inf();
    x0 = active ? x : inf;
    y0 = active ? y : inf;
    x1 = active ? x : -inf;
    y1 = active ? y : -inf;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, d));
        y0 = fmin(y0, __shfl_xor(y0, d));
        x1 = fmax(x1, __shfl_xor(x1, d));
        y1 = fmax(y1, __shfl_xor(y1, d));
    }
}

// Runs (sub-cells) of neighbour cells 4*rd .. 4*rd+3 that can hold a point
// within the cutoff of the box (x0, y0, x1, y1): bit 16*(j-4*rd) + sub.
// Lane l tests neighbour 4*rd + l/16, sub-cell l%16, from the run's bounding
// box shifted by the image offset.
static_assert(kSubLog == 2, "one lane per (neighbour, sub-cell) of 4 neighbours");
__device__ __forceinline__ unsigned long long keep_mask(const double4 *__restrict__ box, int rd,
                                                        int lane, int cbase, int cx, int cy, int nsx,
                                                        int nsy, double Lx, double Ly, double lim,
                                                        double x0, double y0, double x1, double y1)
{
    const int j = 4 * rd + (lane >> 4), sub = lane & 15;
    const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
    const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
    const double zx = (Lx * (qx - wx)) / nsx, zy = (Ly * (qy - wy)) / nsy;
    const double4 r = box[((cbase + wy * nsx + wx) << 4) + sub];
    const double gx = fmax(fmax(r.x + zx - x1, x0 - (r.z + zx)), 0.0);
    const double gy = fmax(fmax(r.y + zy - y1, y0 - (r.w + zy)), 0.0);
    return __ballot(gx * gx + gy * gy < lim);
}

// Bounding box (xmin, ymin, xmax, ymax) of every key run of sorted points;
// empty runs get (+inf, +inf, -inf, -inf), which no box is close to.  One
// wavefront per run (runs are short but clustered points make some long).
__global__ __launch_bounds__(256) void k_run_bbox(const double2 *__restrict__ pos,
                                                  const int *__restrict__ off, int nkeys,
                                                  double4 *__restrict__ box)
{
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= nkeys) return;   // uniform over the wavefront
    const double inf = __builtin_inf();
    double x0 = inf, y0 = inf, x1 = -inf, y1 = -inf;
    for (int i = off[k] + lane; i < off[k + 1]; i += 64) {
        const double2 p = pos[i];
        x0 = fmin(x0, p.x);
        y0 = fmin(y0, p.y);
        x1 = fmax(x1, p.x);
        y1 = fmax(y1, p.y);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, d));
        y0 = fmin(y0, __shfl_xor(y0, d));
        x1 = fmax(x1, __shfl_xor(x1, d));
        y1 = fmax(y1, __shfl_xor(y1, d));
    }
    if (lane == 0) box[k] = make_double4(x0, y0, x1, y1);
}

// ---- the cell-list kernel ------------------------------------------------
// Right-hand sides per pass of the several-density pair kernel (8 or 4).
constexpr int kBatchRB = 8;

template <int KIND, int RB>
__global__ __launch_bounds__(256) void k_real(RealArgs a, const double2 *__restrict__ spos,
                                              const double *__restrict__ qsrc,
                                              const int *__restrict__ soff,
                                              const double4 *__restrict__ sbox)
{
    __shared__ double e1tab[KIND == 0 ? SE2P_E1_NPIECE * SE2P_E1_ROW : 1];
    __shared__ double etab[SE2P_EXP2_N];
    __shared__ double red[kWaves - 1][RB * 2][64];
    const int lane = threadIdx.x & 63;
    // wave index as a scalar so that the source loops and loads stay uniform
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (KIND == 0)
        for (int i = threadIdx.x; i < SE2P_E1_NPIECE * SE2P_E1_ROW; i += 256) e1tab[i] = c_e1_table[i];
    for (int i = threadIdx.x; i < SE2P_EXP2_N; i += blockDim.x) etab[i] = c_exp2_table[i];
    __syncthreads();

    // workgroup -> (chunk, group of RB right-hand sides), groups fastest: a
    // chunk's groups are dispatched together, so the launch does not end on
    // the heavy chunks of the last group (5% on 64 densities)
    const int ngroup = (a.r_end - a.r_begin) / RB;
    const int b = (int)blockIdx.x / ngroup;
    const int grp = (int)blockIdx.x - b * ngroup;
    if (b >= a.chunk_off[a.ncell]) return;   // uniform over the workgroup
    const int cell = a.block_cell[b];
    const RealBox &B = a.boxes.get(a.cell_box ? a.cell_box[cell] : 0);
    const int nsx = B.nsx, nsy = B.nsy, cbase = B.cell_base;
    const double Lx = B.Lx, Ly = B.Ly, cut2 = B.cut2;
    const double xi = B.xi, xi2 = B.xi2, self = B.self;
    const int sub = 1 << (2 * kSubLog);
    const int t0 = a.toff[cell * sub] + 64 * (b - a.chunk_off[cell]);
    const int t1 = a.toff[(cell + 1) * sub];
    const int t = t0 + lane;
    const bool active = t < t1;
    double tx, ty;
    if (active) {
        const double2 p = a.tpos[t];
        tx = p.x;
        ty = p.y;
    } else {
        tx = ty = __builtin_nan("");   // inactive lanes compute garbage, never stored
    }
    const int ncomp = KIND == 0 ? 2 : kDlpComp;
    const int64_t qrow = (int64_t)ncomp * a.nrhs;   // several densities: one source's values
    const int lcell = cell - cbase;
    const int cx = lcell % nsx, cy = lcell / nsx;

    // neighbour sub-cells within reach of this chunk's targets
    double bx0, by0, bx1, by1;
    wave_bbox(active, tx, ty, bx0, by0, bx1, by1);
    const double lim = cut2 * (1.0 + 1e-9);
    const unsigned long long keep0 =
        keep_mask(sbox, 0, lane, cbase, cx, cy, nsx, nsy, Lx, Ly, lim, bx0, by0, bx1, by1);
    const unsigned long long keep1 =
        keep_mask(sbox, 1, lane, cbase, cx, cy, nsx, nsy, Lx, Ly, lim, bx0, by0, bx1, by1);

    for (int r0 = a.r_begin + grp * RB; r0 < a.r_begin + (grp + 1) * RB && r0 + RB <= a.r_end;
         r0 += RB) {
        double acc[RB][2];
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r][0] = acc[r][1] = 0.0;
        const double *qb = qsrc + (int64_t)r0 * ncomp;   // + s * qrow: source s's group
        // a source's position (its densities are read per pair from qb)
        struct Src {
            double x, y;
        };
        auto load = [&](int s) {
            const double2 sp = spos[s];
            return Src{sp.x, sp.y};
        };
        auto pair = [&](const Src &src, int s, double zx, double zy, bool own) {
            const double sx = src.x, sy = src.y;
            double d1, d2;
            if (own) {
                // own cell, no image: both kernels' difference is src - tgt
                // (up to the sign of a zero, which only a skipped r^2 == 0 has)
                d1 = sx - tx;
                d2 = sy - ty;
            } else {
                diff<KIND>(sx, sy, tx, ty, zx, zy, d1, d2);
            }
            double d11, d22;
            const double r2 = se2p_sqsum2(d1, d2, &d11, &d22);
            if (own ? (r2 == 0.0) : !(r2 < cut2)) {
                if (KIND == 0 && own) {   // coincident points: the self term
#pragma unroll
                    for (int r = 0; r < RB; ++r) {
                        acc[r][0] += self * qb[s * qrow + 2 * r];
                        acc[r][1] += self * qb[s * qrow + 2 * r + 1];
                    }
                }
                return;
            }
            if (KIND == 0) slp_pair<RB>(d1, d2, r2, xi, e1tab, etab, qb + s * qrow, 1, acc);
            else dlp_pair<RB>(d1, d2, d11, d22, r2, xi2, etab, qb + s * qrow, 1, acc);
        };

        // sources s0 + w, s0 + w + kWaves, ... < s1
        auto run = [&](int s, int s1, double zx, double zy, bool own) {
#pragma unroll 2
            for (; s < s1; s += kWaves) pair(load(s), s, zx, zy, own);
        };
        // own cell: every pair, no cutoff
        run(soff[cell * sub] + wave, soff[(cell + 1) * sub], 0.0, 0.0, true);
        // eight neighbour cells, periodic image offsets, r^2 < cut2; the kept
        // runs of a cell are visited in order, wave w taking every 4th source
        // of their concatenation
        int cnt = 0;
#pragma unroll 1
        for (int j = 0; j < 8; ++j) {
            unsigned bits = (unsigned)(((j < 4) ? keep0 : keep1) >> (16 * (j & 3))) & 0xffffu;
            if (bits == 0) continue;
            const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
            const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
            const int sb = cbase + wy * nsx + wx;
            const double zx = (Lx * (qx - wx)) / nsx;   // :248-249
            const double zy = (Ly * (qy - wy)) / nsy;
            while (bits) {
                const int ra = __builtin_ctz(bits);
                const int rb = ra + __builtin_ctz(~(bits >> ra));
                bits &= ~((1u << rb) - 1u);
                const int s0 = soff[sb * sub + ra], s1 = soff[sb * sub + rb];
                run(s0 + ((wave - cnt) & (kWaves - 1)), s1, zx, zy, false);
                cnt += s1 - s0;
            }
        }
        // fixed-order reduction of the four partial sums
        if (wave > 0) {
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                red[wave - 1][2 * r][lane] = acc[r][0];
                red[wave - 1][2 * r + 1][lane] = acc[r][1];
            }
        }
        __syncthreads();
        if (wave == 0 && active) {
            const int o = a.tperm[t] - B.t_base;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                double v0 = acc[r][0], v1 = acc[r][1];
#pragma unroll
                for (int w = 0; w < kWaves - 1; ++w) {
                    v0 += red[w][2 * r][lane];
                    v1 += red[w][2 * r + 1][lane];
                }
                double *ur = B.u + (int64_t)(r0 + r) * a.ustride;
                if (KIND == 0) {   // :344-345
                    ur[2 * o] = v0 / (4 * SE2P_PI);
                    ur[2 * o + 1] = v1 / (4 * SE2P_PI);
                } else {           // mex_stokes_dlp_real.cpp:198-199
                    ur[2 * o] = v0 / 4 / SE2P_PI;
                    ur[2 * o + 1] = v1 / 4 / SE2P_PI;
                }
            }
        }
        __syncthreads();
    }
}

// ---- the single-density kernel over windows of a chunk's sources --------
// A chunk's source sequence is its own cell's sources followed by the kept
// runs of its eight neighbour cells in (neighbour, sub-cell) order.  It is
// cut into windows of kWin positions.  In a window, wave w sums the
// positions p = w (mod 4) in order from zero; the four wave sums are added
// in wave order, and the window sums in window order from zero.  That sum
// is the same whether one workgroup walks every window of a chunk or a
// chunk is split into one work item per window (its window sums go to
// slots that k_real_win_final adds), so the split, which balances the
// launch, never changes a result, and neither does the order in which
// work items run.  Work items are numbered in chunk order; persistent
// workgroups take them from eight queues, one per XCD, each a contiguous
// range of items, and drain the other queues when their own is empty.
constexpr int kWin = 256;
constexpr int kQueueStride = 32;   // ints between queue counters (own cache lines)
constexpr int kWinSlots = 16384;   // window slots of split chunks (16 MiB of sums)
constexpr int kWinBlocksPerCU = 8;
// When other work runs beside the pair kernel (shares_gpu: a concurrent
// group, se2p_group_begin, or a driver call's two sums), its persistent
// workgroups leave a quarter of each CU's wave slots to that work instead of
// taking them all until the last item.
constexpr int kWinBlocksPerCUGroup = 6;

struct WinPlan {
    unsigned long long *keep;   // [2 max_chunks] kept neighbour runs (keep_mask bits)
    int *npos;                  // [max_chunks] sequence length
    int2 *where;                // [max_chunks] (first target, cell)
    int *item_off;              // [max_chunks] first work item of each chunk
    int *slot;                  // [max_chunks] first window slot of a split chunk, -1 if unsplit
    int *item_chunk;            // [max_chunks + slots] work item -> chunk
    int4 *split;                // [max_chunks] split chunks: (slot, windows, first target, cell)
    int *meta;                  // [0] work items, [1] split threshold, [2] split chunks,
                                // queue x at kQueueStride (x + 1)
    double *part;               // [slots][2][64] window sums of split chunks
    int slots;
};

__device__ __forceinline__ int win_count(int n) { return n > kWin ? (n + kWin - 1) / kWin : 1; }

// Sequence length and kept runs of chunk c (one wavefront).
__device__ __forceinline__ void win_plan_chunk(const RealArgs &a, const int *__restrict__ soff,
                                               const double4 *__restrict__ sbox, const WinPlan &w,
                                               int c)
{
    const int lane = threadIdx.x & 63;
    const int cell = a.block_cell[c];
    const RealBox &B = a.boxes.get(a.cell_box ? a.cell_box[cell] : 0);
    const int nsx = B.nsx, nsy = B.nsy, cbase = B.cell_base;
    const int sub = 1 << (2 * kSubLog);
    const int t0 = a.toff[cell * sub] + 64 * (c - a.chunk_off[cell]);
    const bool active = t0 + lane < a.toff[(cell + 1) * sub];
    double tx = 0.0, ty = 0.0;
    if (active) {
        const double2 p = a.tpos[t0 + lane];
        tx = p.x;
        ty = p.y;
    }
    const int lcell = cell - cbase;
    const int cx = lcell % nsx, cy = lcell / nsx;
    double bx0, by0, bx1, by1;
    wave_bbox(active, tx, ty, bx0, by0, bx1, by1);
    const double lim = B.cut2 * (1.0 + 1e-9);
    unsigned long long keep[2];
    int len = 0;
#pragma unroll
    for (int rd = 0; rd < 2; ++rd) {
        keep[rd] = keep_mask(sbox, rd, lane, cbase, cx, cy, nsx, nsy, B.Lx, B.Ly, lim, bx0, by0,
                             bx1, by1);
        const int j = 4 * rd + (lane >> 4);
        const int wx = (cx + nb_dx(j) + nsx) % nsx, wy = (cy + nb_dy(j) + nsy) % nsy;
        const int k = ((cbase + wy * nsx + wx) << 4) + (lane & 15);
        if ((keep[rd] >> lane) & 1) len += soff[k + 1] - soff[k];
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) len += __shfl_xor(len, d);
    if (lane == 0) {
        w.keep[2 * c] = keep[0];
        w.keep[2 * c + 1] = keep[1];
        w.npos[c] = soff[(cell + 1) * sub] - soff[cell * sub] + len;
        w.where[c] = make_int2(t0, cell);
    }
}

// Work items: a chunk with at least `theta` windows is split into one item
// per window, the others are one item each.  theta is the smallest of
// 2, 4, ..., 64 whose split chunks fit the window slots (none split if even
// 64 does not).  Split chunks are also listed, with what k_real_win_final
// needs, for that kernel.  One workgroup of kWinScanWaves wavefronts;
// wavefront v owns a contiguous range of chunks and walks it 64 chunks
// (one per lane) at a time, so every load is coalesced.
constexpr int kWinScanWaves = 16;
constexpr int kWinTheta = 6;   // candidate thresholds 2 << i
__global__ __launch_bounds__(64 * kWinScanWaves) void k_real_win_scan(const int *__restrict__ chunk_off,
                                                                      int ncell, WinPlan w)
{
    __shared__ unsigned long long wtot[kWinScanWaves][kWinTheta];
    __shared__ int wcnt[kWinScanWaves][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nch = chunk_off[ncell];
    const int range = (nch + kWinScanWaves - 1) / kWinScanWaves;
    const int r0 = min(wave * range, nch), r1 = min(r0 + range, nch);
    // windows of the chunks with at least 2 << q windows
    unsigned long long sum[kWinTheta] = {};
#pragma unroll 4
    for (int c = r0 + lane; c < r1; c += 64) {
        const int nw = win_count(w.npos[c]);
#pragma unroll
        for (int q = 0; q < kWinTheta; ++q) sum[q] += nw >= (2 << q) ? nw : 0;
    }
#pragma unroll
    for (int q = 0; q < kWinTheta; ++q) {
        for (int d = 32; d > 0; d >>= 1) sum[q] += __shfl_xor(sum[q], d);
        if (lane == 0) wtot[wave][q] = sum[q];
    }
    __syncthreads();
    int theta = 0x7fffffff;
    for (int q = kWinTheta - 1; q >= 0; --q) {
        unsigned long long t = 0;
        for (int v = 0; v < kWinScanWaves; ++v) t += wtot[v][q];
        if (t <= (unsigned long long)w.slots) theta = 2 << q;
    }
    // per chunk: work items, window slots, split-list entries
    auto counts = [&](int nw, int &ni, int &ns, int &nl) {
        const bool split = nw >= theta;
        ni = split ? nw : 1;
        ns = split ? nw : 0;
        nl = split ? 1 : 0;
    };
    int ti = 0, ts = 0, tl = 0;
#pragma unroll 4
    for (int c = r0 + lane; c < r1; c += 64) {
        int ni, ns, nl;
        counts(win_count(w.npos[c]), ni, ns, nl);
        ti += ni;
        ts += ns;
        tl += nl;
    }
    for (int d = 32; d > 0; d >>= 1) {
        ti += __shfl_xor(ti, d);
        ts += __shfl_xor(ts, d);
        tl += __shfl_xor(tl, d);
    }
    if (lane == 0) {
        wcnt[wave][0] = ti;
        wcnt[wave][1] = ts;
        wcnt[wave][2] = tl;
    }
    __syncthreads();
    int ci = 0, cs = 0, cl = 0;   // this wavefront's first item, slot, split-list entry
    for (int v = 0; v < wave; ++v) {
        ci += wcnt[v][0];
        cs += wcnt[v][1];
        cl += wcnt[v][2];
    }
    for (int base = r0; base < r1; base += 64) {
        const int c = base + lane;
        int nw = 0, ni = 0, ns = 0, nl = 0;
        if (c < r1) {
            nw = win_count(w.npos[c]);
            counts(nw, ni, ns, nl);
        }
        // inclusive scans across the wavefront
        int si = ni, ss = ns, sl = nl;
        for (int d = 1; d < 64; d <<= 1) {
            const int a = __shfl_up(si, d), b = __shfl_up(ss, d), e = __shfl_up(sl, d);
            if (lane >= d) {
                si += a;
                ss += b;
                sl += e;
            }
        }
        if (c < r1) {
            const int io = ci + si - ni, so = cs + ss - ns, lo = cl + sl - nl;
            w.item_off[c] = io;
            w.slot[c] = ns ? so : -1;
            for (int k = 0; k < ni; ++k) w.item_chunk[io + k] = c;
            if (ns) {
                const int2 at = w.where[c];
                w.split[lo] = make_int4(so, nw, at.x, at.y);
            }
        }
        ci += __shfl(si, 63);
        cs += __shfl(ss, 63);
        cl += __shfl(sl, 63);
    }
    if (threadIdx.x == 0) {
        int t0 = 0, t2 = 0;
        for (int v = 0; v < kWinScanWaves; ++v) {
            t0 += wcnt[v][0];
            t2 += wcnt[v][2];
        }
        w.meta[0] = t0;
        w.meta[1] = theta;
        w.meta[2] = t2;
    }
    if (threadIdx.x < 8) w.meta[kQueueStride * (threadIdx.x + 1)] = 0;
}

// Sequence length and kept runs of every chunk: one wavefront per chunk.
__global__ __launch_bounds__(256) void k_real_win_plan(RealArgs a, const int *__restrict__ soff,
                                                       const double4 *__restrict__ sbox, WinPlan w)
{
    const int c = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (c < a.chunk_off[a.ncell]) win_plan_chunk(a, soff, sbox, w, c);   // uniform over the wavefront
}

// The stresslet fits 64 VGPRs, eight waves per SIMD (2% faster than seven);
// the Stokeslet 96, five waves (3% faster than its unbounded 99 and four).
template <int KIND>
__global__ __launch_bounds__(256, KIND == 1 ? 8 : 5) void k_real_win(RealArgs a, const double *__restrict__ srec,
                                                  const int *__restrict__ soff, WinPlan w)
{
    __shared__ double e1tab[KIND == 0 ? SE2P_E1_NPIECE * SE2P_E1_ROW : 1];
    __shared__ double etab[SE2P_EXP2_N];
    __shared__ double red[2][kWaves - 1][2][64];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (KIND == 0)
        for (int i = threadIdx.x; i < SE2P_E1_NPIECE * SE2P_E1_ROW; i += 256) e1tab[i] = c_e1_table[i];
    for (int i = threadIdx.x; i < SE2P_EXP2_N; i += blockDim.x) etab[i] = c_exp2_table[i];
    const int total = w.meta[0];
    const int per = (total + 7) >> 3;
    const int xcd = blockIdx.x & 7;
    const int sub = 1 << (2 * kSubLog);
    int qi = 0;   // thread 0: queues found empty so far
    for (;;) {
        if (threadIdx.x == 0) {
            int item = -1;
            for (; qi < 8; ++qi) {
                const int x = (xcd + qi) & 7;
                const int lo = x * per, nq = min(per, total - lo);
                if (nq <= 0) continue;
                const int it = atomicAdd(&w.meta[kQueueStride * (x + 1)], 1);
                if (it < nq) {
                    item = lo + it;
                    break;
                }
            }
            s_item = item;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item < 0) break;   // uniform over the workgroup

        const int c = w.item_chunk[item];
        const int sl = w.slot[c];
        const int cell = a.block_cell[c];
        const RealBox &B = a.boxes.get(a.cell_box ? a.cell_box[cell] : 0);
        const int nsx = B.nsx, nsy = B.nsy, cbase = B.cell_base;
        const double Lx = B.Lx, Ly = B.Ly, cut2 = B.cut2;
        const double xi = B.xi, xi2 = B.xi2, self = B.self;
        const int t = a.toff[cell * sub] + 64 * (c - a.chunk_off[cell]) + lane;
        const bool active = t < a.toff[(cell + 1) * sub];
        double tx, ty;
        if (active) {
            const double2 p = a.tpos[t];
            tx = p.x;
            ty = p.y;
        } else {
            tx = ty = __builtin_nan("");   // inactive lanes compute garbage, never stored
        }
        const int lcell = cell - cbase;
        const int cx = lcell % nsx, cy = lcell / nsx;
        const int n = w.npos[c];
        const int kb = sl >= 0 ? item - w.item_off[c] : 0;
        const int ke = sl >= 0 ? kb + 1 : win_count(n);
        const unsigned long long keep0 = w.keep[2 * c], keep1 = w.keep[2 * c + 1];

        double acc0 = 0.0, acc1 = 0.0;
        struct Src {
            double x, y, q[6];
        };
        auto load = [&](int s) {
            Src o;
            const double *rp = srec + (int64_t)s * SrcRec<KIND>::n;
            const double4 v0 = *(const double4 *)rp;
            o.x = v0.x;
            o.y = v0.y;
            o.q[0] = v0.z;
            o.q[1] = v0.w;
            if (KIND == 1) {
                const double4 v1 = *(const double4 *)(rp + 4);
                o.q[2] = v1.x;
                o.q[3] = v1.y;
                o.q[4] = v1.z;
                o.q[5] = v1.w;
            }
            return o;
        };
        // own: the own cell (no cutoff); img: a neighbour seen through a
        // periodic image (zx, zy not both zero)
        auto pair = [&](const Src &src, double zx, double zy, bool own, bool img) {
            const double *q0 = src.q;
            double d1, d2;
            if (img) {
                diff<KIND>(src.x, src.y, tx, ty, zx, zy, d1, d2);
            } else {
                // no image offset: both kernels' difference is src - tgt (up
                // to the sign of a zero, which adding a zero offset can flip
                // and which no sum sees)
                d1 = src.x - tx;
                d2 = src.y - ty;
            }
            double d11, d22;
            const double r2 = se2p_sqsum2(d1, d2, &d11, &d22);
            if (own ? (r2 == 0.0) : !(r2 < cut2)) {
                if (KIND == 0 && own) {   // coincident points: the self term
                    acc0 += self * q0[0];
                    acc1 += self * q0[1];
                }
                return;
            }
            if (KIND == 0) {
                double hE, g;
                slp_factors(r2, xi, e1tab, etab, hE, g);
                slp_add(d1, d2, hE, g, q0[0], q0[1], acc0, acc1);
            } else {
                // record values carry -4 and 2 xi^2:
                //   u_i += e2 (H d_i + l'_i),  H = (1 + x)/r^4 (d.S'd),
                // with (1 + x)/r^4 = (1/r^2)(1/r^2 + xi^2)
                const double x = xi2 * r2;
                const double e2 = se2p_exp_neg(x, etab);
                const double ir2 = rcp_nr(r2);
                const double H = (ir2 * (ir2 + xi2)) * (d11 * q0[0] + (d1 * d2) * q0[1] + d22 * q0[2]);
                acc0 = fma(e2, fma(H, d1, d1 * q0[3] + d2 * q0[5]), acc0);
                acc1 = fma(e2, fma(H, d2, d1 * q0[5] + d2 * q0[4]), acc1);
            }
        };
        // sources s, s + kWaves, ... < s1, two records loaded together
        auto run = [&](int s, int s1, double zx, double zy, bool own, bool img) {
            for (; s + kWaves < s1; s += 2 * kWaves) {
                const Src a0 = load(s), a1 = load(s + kWaves);
                pair(a0, zx, zy, own, img);
                pair(a1, zx, zy, own, img);
            }
            if (s < s1) pair(load(s), zx, zy, own, img);
        };

        // the walk: current run [s0, s1) at sequence position pos; j = -1 is
        // the own cell, j = 0..7 the neighbours (mex_stokes_slp_real.cpp:61-62)
        int j = -1, s0 = soff[cell * sub], s1 = soff[(cell + 1) * sub], pos = 0, sb = 0;
        unsigned bits = 0;
        double zx = 0.0, zy = 0.0;
        bool more = true;
        auto advance = [&]() {
            pos += s1 - s0;
            while (bits == 0) {
                if (++j == 8) {
                    more = false;
                    return;
                }
                bits = (unsigned)(((j < 4) ? keep0 : keep1) >> (16 * (j & 3))) & 0xffffu;
                if (bits) {
                    const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
                    const int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
                    sb = cbase + wy * nsx + wx;
                    zx = (Lx * (qx - wx)) / nsx;   // :248-249
                    zy = (Ly * (qy - wy)) / nsy;
                }
            }
            const int ra = __builtin_ctz(bits);
            const int rb = ra + __builtin_ctz(~(bits >> ra));
            bits &= ~((1u << rb) - 1u);
            s0 = soff[sb * sub + ra];
            s1 = soff[sb * sub + rb];
        };
        double tot0 = 0.0, tot1 = 0.0;
        for (int k = kb; k < ke; ++k) {
            const int wa = k * kWin, wb = min(wa + kWin, n);
            acc0 = acc1 = 0.0;
            while (more) {
                const int pe = pos + (s1 - s0);
                const int pa = max(pos, wa), pb = min(pe, wb);
                if (pa < pb) {
                    const int sa = s0 + (pa - pos) + ((wave - pa) & (kWaves - 1));
                    const int se = s0 + (pb - pos);
                    if (j < 0) run(sa, se, 0.0, 0.0, true, false);
                    else if (zx == 0.0 && zy == 0.0) run(sa, se, 0.0, 0.0, false, false);
                    else run(sa, se, zx, zy, false, true);
                }
                if (pe > wb) break;   // the run goes on in the next window
                advance();
            }
            // the window's sum: wave sums added in wave order
            if (wave > 0) {
                red[k & 1][wave - 1][0][lane] = acc0;
                red[k & 1][wave - 1][1][lane] = acc1;
            }
            __syncthreads();
            if (wave == 0) {
                double v0 = acc0, v1 = acc1;
#pragma unroll
                for (int q = 0; q < kWaves - 1; ++q) {
                    v0 += red[k & 1][q][0][lane];
                    v1 += red[k & 1][q][1][lane];
                }
                if (sl >= 0) {
                    // the lane index recomputed here (mbcnt): a lane offset
                    // kept live across the item loop spilled
                    const int lo = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
                    double *pp = w.part + (int64_t)(sl + k) * 128;
                    pp[lo] = v0;
                    pp[64 + lo] = v1;
                } else {
                    tot0 += v0;
                    tot1 += v1;
                }
            }
        }
        if (sl < 0 && wave == 0 && active) {
            const int o = a.tperm[t] - B.t_base;
            if (KIND == 0) {   // :344-345
                B.u[2 * o] = tot0 / (4 * SE2P_PI);
                B.u[2 * o + 1] = tot1 / (4 * SE2P_PI);
            } else {           // mex_stokes_dlp_real.cpp:198-199
                B.u[2 * o] = tot0 / 4 / SE2P_PI;
                B.u[2 * o + 1] = tot1 / 4 / SE2P_PI;
            }
        }
    }
}

// Split chunks: window sums added in window order from zero, then scaled
// and scattered like k_real_win's own results.  One wavefront per split
// chunk, the wavefronts of the launch striding over the list.
constexpr int kWinFinalBlocks = 256;
template <int KIND>
__global__ __launch_bounds__(256) void k_real_win_final(RealArgs a, WinPlan w)
{
    const int lane = threadIdx.x & 63;
    const int nsplit = w.meta[2];
    const int sub = 1 << (2 * kSubLog);
    for (int i = blockIdx.x * kWaves + (threadIdx.x >> 6); i < nsplit; i += gridDim.x * kWaves) {
        const int4 r = w.split[i];   // (slot, windows, first target, cell)
        const int t = r.z + lane;
        if (t >= a.toff[(r.w + 1) * sub]) continue;
        const double *pp = w.part + (int64_t)r.x * 128 + lane;
        const int nw = r.y;
        double v0 = 0.0, v1 = 0.0;
        int k = 0;
        for (; k + 8 <= nw; k += 8) {   // eight windows' loads in flight, added in order
            double x0[8], x1[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                x0[q] = pp[(k + q) * 128];
                x1[q] = pp[(k + q) * 128 + 64];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                v0 += x0[q];
                v1 += x1[q];
            }
        }
        for (; k < nw; ++k) {
            v0 += pp[k * 128];
            v1 += pp[k * 128 + 64];
        }
        const RealBox &B = a.boxes.get(a.cell_box ? a.cell_box[r.w] : 0);
        const int o = a.tperm[t] - B.t_base;
        if (KIND == 0) {
            B.u[2 * o] = v0 / (4 * SE2P_PI);
            B.u[2 * o + 1] = v1 / (4 * SE2P_PI);
        } else {
            B.u[2 * o] = v0 / 4 / SE2P_PI;
            B.u[2 * o + 1] = v1 / 4 / SE2P_PI;
        }
    }
}

// Pair statistics: one wave per chunk, counts pairs evaluated / rejected
// (every neighbour-cell pair the reference examines).
__global__ __launch_bounds__(64) void k_real_count(RealArgs a)
{
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    if (b >= a.chunk_off[a.ncell]) return;
    const int cell = a.block_cell[b];
    const RealBox &B = a.boxes.get(a.cell_box ? a.cell_box[cell] : 0);
    const int sub = 1 << (2 * kSubLog);
    const int t = a.toff[cell * sub] + 64 * (b - a.chunk_off[cell]) + lane;
    if (t >= a.toff[(cell + 1) * sub]) return;
    const double2 p = a.tpos[t];
    unsigned long long n_in = 0, n_out = 0;
    const int lcell = cell - B.cell_base;
    const int cx = lcell % B.nsx, cy = lcell / B.nsx;
    for (int s = a.soff[cell * sub]; s < a.soff[(cell + 1) * sub]; ++s) {
        double d1, d2;
        diff<1>(a.spos[s].x, a.spos[s].y, p.x, p.y, 0.0, 0.0, d1, d2);
        n_in += se2p_sqsum(d1, d2) != 0.0 ? 1 : 0;
    }
    for (int j = 0; j < 8; ++j) {
        const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
        const int wx = (qx + B.nsx) % B.nsx, wy = (qy + B.nsy) % B.nsy;
        const int sb = B.cell_base + wy * B.nsx + wx;
        const double zx = (B.Lx * (qx - wx)) / B.nsx;
        const double zy = (B.Ly * (qy - wy)) / B.nsy;
        for (int s = a.soff[sb * sub]; s < a.soff[(sb + 1) * sub]; ++s) {
            double d1, d2;
            diff<1>(a.spos[s].x, a.spos[s].y, p.x, p.y, zx, zy, d1, d2);
            if (se2p_sqsum(d1, d2) < B.cut2) ++n_in;
            else ++n_out;
        }
    }
    atomicAdd(&a.pair_counts[0], n_in);   // integer sums: order-independent
    atomicAdd(&a.pair_counts[1], n_out);
}

// ---- adjoint ---------------------------------------------------------------
// With u the output and L = sum_t du_t . u_t, a pair (target t, source s, image
// offset z) contributes phi(d) / (4 pi) to L, d = x_s - x_t + z, phi = du_t . K(d) q_s.
// Then dL/dx_s += grad phi / (4 pi), dL/dx_t -= grad phi / (4 pi), and the
// density gradient is the transposed pair sum.  The cell and cutoff
// decisions are those of the forward sum (taken as fixed).
//
// Stokeslet: phi = a . G(d) b, a = du_t, b = f_s, G = E1/2 I + e (d d^T/r^2 - I),
//   grad phi = e { [-(s + 2 p q/r^2)/r^2 - 2 xi^2 (p q/r^2 - s)] d + (q a + p b)/r^2 }
//   with p = d.a, q = d.b, s = a.b, e = exp(-xi^2 r^2); the E1 term gives
//   -e d/r^2 (dE1/dx = -e^{-x}/x).
__device__ __forceinline__ void slp_grad(double d1, double d2, double xi2, double e, double ir2,
                                         double a1, double a2, double b1, double b2, double &g1,
                                         double &g2)
{
    const double p = d1 * a1 + d2 * a2, q = d1 * b1 + d2 * b2, s = a1 * b1 + a2 * b2;
    const double pq = p * q * ir2;
    const double ei = e * ir2;
    const double cd = -ei * (s + 2.0 * pq) - 2.0 * xi2 * e * (pq - s);
    g1 += cd * d1 + ei * (q * a1 + p * b1);
    g2 += cd * d2 + ei * (q * a2 + p * b2);
}

// Stresslet: phi = e2 [A p c + B (2 a.S d + p tr S)] with S the symmetric
// density matrix [[q1, q2/2], [q2/2, q3]], A = -4 (1 + x)/r^4, B = 2 xi^2,
// x = xi^2 r^2, p = d.a, c = d.S d;
//   grad phi = [-2 xi^2 phi + e2 A' p c] d + e2 [A (c a + 2 p S d) + B (2 S a + tr S a)],
//   A' d = grad A, A' = 8 (2 + x)/r^6.
__device__ __forceinline__ void dlp_grad(double d1, double d2, double r2, double xi2, double e2,
                                         double ir2, double a1, double a2, double q1, double q2,
                                         double q3, double &g1, double &g2)
{
    const double x = xi2 * r2;
    const double ir4 = ir2 * ir2;
    const double A = -4.0 * (1.0 + x) * ir4;
    const double Ap = 8.0 * (2.0 + x) * ir4 * ir2;
    const double B = 2.0 * xi2;
    const double hq = 0.5 * q2;
    const double sd1 = q1 * d1 + hq * d2, sd2 = hq * d1 + q3 * d2;
    const double sa1 = q1 * a1 + hq * a2, sa2 = hq * a1 + q3 * a2;
    const double p = d1 * a1 + d2 * a2;
    const double c = d1 * sd1 + d2 * sd2;
    const double tr = q1 + q3;
    const double phi = e2 * (A * p * c + B * (2.0 * (a1 * sd1 + a2 * sd2) + p * tr));
    const double cd = -2.0 * xi2 * phi + e2 * Ap * p * c;
    g1 += cd * d1 + e2 * (A * (c * a1 + 2.0 * p * sd1) + B * (2.0 * sa1 + tr * a1));
    g2 += cd * d2 + e2 * (A * (c * a2 + 2.0 * p * sd2) + B * (2.0 * sa2 + tr * a2));
}

// Stresslet density gradient: m_j += sum_i a_i T_i(j), for the three merged
// densities (S11, S12 + S21, S22).
__device__ __forceinline__ void dlp_dens(double d1, double d2, double r2, double xi2, double e2,
                                         double ir2, double a1, double a2, double &m1, double &m2,
                                         double &m3)
{
    const double fb = -4.0 * (1.0 + xi2 * r2) * ir2 * ir2;
    const double pre = 2.0 * xi2;
    const double d11 = d1 * d1, d12 = d1 * d2, d22 = d2 * d2;
    const double t111 = e2 * (d11 * d1 * fb + pre * 3.0 * d1);
    const double t112 = e2 * (d11 * d2 * fb + pre * d2);
    const double t122 = e2 * (d12 * d2 * fb + pre * d1);
    const double t222 = e2 * (d22 * d2 * fb + pre * 3.0 * d2);
    m1 += a1 * t111 + a2 * t112;
    m2 += a1 * t112 + a2 * t122;
    m3 += a1 * t122 + a2 * t222;
}

struct RealAdjArgs {
    const double2 *lpos;   // lane points in (cell, sub-cell) order
    const int *lperm;      // that order -> caller's index
    const int *loff;       // [ncell*16+1] run offsets of the lane points
    const int *chunk_off;  // [ncell+1] exclusive scan of ceil(lane points/64)
    const int *block_cell; // workgroup -> cell
    const double *ldata;   // lane data in lane order, [ncomp][nl]
    int64_t nl, nu;        // lane / uniform point counts (data strides)
    int nsx, nsy, ncell;
    double Lx, Ly, xi, xi2, cut2, self;
    const double *f, *nv;  // caller-order densities and normals (stresslet sources)
    double *dpos, *df, *dn;   // caller order; may be null
};

// SIDE 0: lanes are targets (lane data du), uniform points are sources
//   (uniform data: the densities); writes dL/dptar.
// SIDE 1: lanes are sources (lane data: the densities), uniform points are
//   targets (uniform data du); writes dL/dpsrc, dL/df, dL/dn.  The image offset
//   from a source cell to a target cell is the negated offset of the forward
//   visit, so d = x_s - (x_t + z') (SLP) or (x_s - x_t) - z' (DLP) repeats the
//   forward's rounding bit for bit and every cutoff test agrees.
template <int KIND, int SIDE>
__global__ __launch_bounds__(256) void k_real_adj(RealAdjArgs a, const double2 *__restrict__ upos,
                                                  const double *__restrict__ udata,
                                                  const int *__restrict__ uoff,
                                                  const double4 *__restrict__ ubox)
{
    // accumulators: SIDE 0 [grad]; SIDE 1 [grad, density terms]
    constexpr int NS = KIND == 0 ? 4 : 5;
    constexpr int NV = SIDE == 0 ? 2 : NS;
    constexpr int SO = 0;   // offset of the source-side accumulators
    __shared__ double e1tab[KIND == 0 && SIDE >= 1 ? SE2P_E1_NPIECE * SE2P_E1_ROW : 1];
    __shared__ double etab[SE2P_EXP2_N];
    __shared__ double red[kWaves - 1][NV][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (KIND == 0 && SIDE >= 1)
        for (int i = threadIdx.x; i < SE2P_E1_NPIECE * SE2P_E1_ROW; i += 256) e1tab[i] = c_e1_table[i];
    for (int i = threadIdx.x; i < SE2P_EXP2_N; i += blockDim.x) etab[i] = c_exp2_table[i];
    __syncthreads();

    const int b = xcd_chunk(a.chunk_off[a.ncell]);
    if (b < 0) return;   // uniform over the workgroup
    const int cell = a.block_cell[b];
    const int sub = 1 << (2 * kSubLog);
    const int l0 = a.loff[cell * sub] + 64 * (b - a.chunk_off[cell]);
    const int l1 = a.loff[(cell + 1) * sub];
    const int l = l0 + lane;
    const bool active = l < l1;
    double lx, ly;
    // lane data: SIDE 0 du; SIDE 1 densities
    double ld[3] = {0.0, 0.0, 0.0};
    constexpr int NQ = KIND == 0 ? 2 : 3;
    constexpr int LC = SIDE == 0 ? 2 : NQ;
    if (active) {
        const double2 p = a.lpos[l];
        lx = p.x;
        ly = p.y;
#pragma unroll
        for (int c = 0; c < LC; ++c)
            ld[c] = a.ldata[c * a.nl + l];
    } else {
        lx = ly = __builtin_nan("");   // inactive lanes compute garbage, never stored
    }
    const int cx = cell % a.nsx, cy = cell / a.nsx;
    const int64_t us = a.nu;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;

    // one pair: d with the forward's rounding, then the side's contributions
    auto pair = [&](double ux, double uy, double zx, double zy, int u, bool own) {
        double d1, d2;
        if (SIDE == 0) {
            diff<KIND>(ux, uy, lx, ly, zx, zy, d1, d2);
        } else if (KIND == 0) {
            d1 = lx - (ux + zx);
            d2 = ly - (uy + zy);
        } else {
            d1 = (lx - ux) - zx;
            d2 = (ly - uy) - zy;
        }
        const double r2 = se2p_sqsum(d1, d2);
        if (own ? (r2 == 0.0) : !(r2 < a.cut2)) {
            if (KIND == 0 && SIDE == 1 && own) {   // coincident: self term, no gradient
                acc[2] += a.self * udata[u];
                acc[3] += a.self * udata[us + u];
            }
            return;
        }
        const double x = r2 * a.xi2;
        const double e = se2p_exp_neg(x, etab);
        const double ir2 = rcp_nr(r2);
        if (SIDE == 0) {
            // lane: du_t; uniform: the source densities
            if (KIND == 0)
                slp_grad(d1, d2, a.xi2, e, ir2, ld[0], ld[1], udata[u], udata[us + u], acc[0], acc[1]);
            else
                dlp_grad(d1, d2, r2, a.xi2, e, ir2, ld[0], ld[1], udata[u], udata[us + u],
                         udata[2 * us + u], acc[0], acc[1]);
        } else {
            // lane: the source densities; uniform: du_t
            const double a1 = udata[u], a2 = udata[us + u];
            if (KIND == 0) {
                slp_grad(d1, d2, a.xi2, e, ir2, a1, a2, ld[0], ld[1], acc[0], acc[1]);
                // G(d) du_t, arranged as the forward pair
                const double E = e1_lds(x, e, e1tab);
                const double hE = 0.5 * E, g = e * ir2;
                const double c = d1 * a2 - d2 * a1;
                acc[2] += hE * a1 + g * d2 * c;
                acc[3] += hE * a2 - g * d1 * c;
            } else {
                dlp_grad(d1, d2, r2, a.xi2, e, ir2, a1, a2, ld[0], ld[1], ld[2], acc[0], acc[1]);
                dlp_dens(d1, d2, r2, a.xi2, e, ir2, a1, a2, acc[2], acc[3], acc[4]);
            }
        }
    };

    {   // own cell: every pair, no cutoff
        const int u1 = uoff[(cell + 1) * sub];
#pragma unroll 2
        for (int u = uoff[cell * sub] + wave; u < u1; u += kWaves) {
            const double2 up = upos[u];
            pair(up.x, up.y, 0.0, 0.0, u, true);
        }
    }
    // neighbour cells: only the uniform points' sub-cells within reach
    double bx0, by0, bx1, by1;
    wave_bbox(active, lx, ly, bx0, by0, bx1, by1);
    const double lim = a.cut2 * (1.0 + 1e-9);
    const unsigned long long keep0 =
        keep_mask(ubox, 0, lane, 0, cx, cy, a.nsx, a.nsy, a.Lx, a.Ly, lim, bx0, by0, bx1, by1);
    const unsigned long long keep1 =
        keep_mask(ubox, 1, lane, 0, cx, cy, a.nsx, a.nsy, a.Lx, a.Ly, lim, bx0, by0, bx1, by1);
    int cnt = 0;
#pragma unroll 1
    for (int j = 0; j < 8; ++j) {
        unsigned bits = (unsigned)(((j < 4) ? keep0 : keep1) >> (16 * (j & 3))) & 0xffffu;
        if (bits == 0) continue;
        const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
        const int wx = (qx + a.nsx) % a.nsx, wy = (qy + a.nsy) % a.nsy;
        const int ub = wy * a.nsx + wx;
        const double zx = (a.Lx * (qx - wx)) / a.nsx;
        const double zy = (a.Ly * (qy - wy)) / a.nsy;
        while (bits) {
            const int ra = __builtin_ctz(bits);
            const int rb = ra + __builtin_ctz(~(bits >> ra));
            bits &= ~((1u << rb) - 1u);
            const int u0 = uoff[ub * sub + ra], u1 = uoff[ub * sub + rb];
#pragma unroll 2
            for (int u = u0 + ((wave - cnt) & (kWaves - 1)); u < u1; u += kWaves) {
                const double2 up = upos[u];
                pair(up.x, up.y, zx, zy, u, false);
            }
            cnt += u1 - u0;
        }
    }
    if (wave > 0)
#pragma unroll
        for (int v = 0; v < NV; ++v) red[wave - 1][v][lane] = acc[v];
    __syncthreads();
    if (wave != 0 || !active) return;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int w = 0; w < kWaves - 1; ++w) acc[v] += red[w][v][lane];
    const int o = a.lperm[l];
    const double s4 = 1.0 / (4 * SE2P_PI);
    if (SIDE == 0) {
        a.dpos[2 * o] = -acc[0] * s4;
        a.dpos[2 * o + 1] = -acc[1] * s4;
        return;
    }
    if (a.dpos) {
        a.dpos[2 * o] = acc[SO] * s4;
        a.dpos[2 * o + 1] = acc[SO + 1] * s4;
    }
    if (KIND == 0) {
        if (a.df) {
            a.df[2 * o] = acc[SO + 2] * s4;
            a.df[2 * o + 1] = acc[SO + 3] * s4;
        }
    } else {
        // m = dL/d(S11, S12 + S21, S22), S = f (x) n
        const double m1 = acc[SO + 2] * s4, m2 = acc[SO + 3] * s4, m3 = acc[SO + 4] * s4;
        const double f1 = a.f[2 * o], f2 = a.f[2 * o + 1];
        const double n1 = a.nv[2 * o], n2 = a.nv[2 * o + 1];
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

// du in target (cell) order, [2][n]
// ---- adjoint, targets are the sources -------------------------------------
// One sorted point set (the forward's plan: source records, cell lists, chunk
// plans, run boxes).  Lane i and uniform point j give two ordered pairs: i as
// target of source j (the target side, dL/dptar) and i as source of target j
// (the source side, dL/dpsrc and the density gradient).  Each uses the
// difference vector and inclusion test the forward sums use.
//
// Stresslet (k_real_adj_win): the lane-as-source difference is exactly the
// negated lane-as-target one (IEEE subtraction is antisymmetric and the image
// offsets of the two visits are negatives), so one r^2 and one test serve
// both.  With
// t = e2 T (the three-index stresslet tensor, fully symmetric; components
// t111, t112, t122, t222) and the merged densities q = (S11, S12+S21, S22),
//   phi(a, q) = sum_k w_k(a, q) t_k,
//   w = (a1 q1, a1 q2 + a2 q1, a1 q3 + a2 q2, a2 q3),
// so grad phi = sum_k w_k D_k with D_lk = d_l t_k: the 2 x 4 tensor D is
// computed once per pair and contracted with the target side's weights
// (lane du, uniform densities) and the source side's (uniform du, lane
// densities); grad t is even in d, t is odd.  The lane's factor of each
// weight is constant over the pass, so it is applied once at the end.  With t = alpha P3(d) + beta P1(d),
// alpha = -4 e2 (1 + x)/r^4, beta = 2 xi^2 e2, x = xi^2 r^2:
//   D_lk = d_l [g1 P3_k + g2 P1_k] + alpha dP3_k/dd_l + beta dP1_k/dd_l,
//   g1 = 8 e2 (x^2 + 2 x + 2)/r^6, g2 = -2 xi^2 beta,
//   P3 = (d1^3, d1^2 d2, d1 d2^2, d2^3), P1 = (3 d1, d2, d1, 3 d2).
struct RealAdjSameArgs {
    const double *rec;       // source records in (cell, sub-cell) order
    const double2 *du;       // du in the same order
    const int *off, *perm;   // run offsets, order -> caller index
    const int *chunk_off, *block_cell;
    const double4 *box;      // run bounding boxes
    int ncell, nsx, nsy;
    double Lx, Ly, xi2, cut2, self;
    const double *f, *nv;    // caller-order densities and normals (stresslet df, dn)
    double *dpsrc, *dptar, *df, *dn;   // caller order; may be null
};

// Stokeslet: each wavefront owns 16 of the chunk's 64 targets and its four
// 16-lane rows take four consecutive sources per step (lane = target + 16 x
// row); a wavefront's reach is then a quarter of the chunk, so fewer of its
// steps carry no pair in range (5% faster at N = 4e6 than 64 targets per
// wavefront).  The rows' sums are added across lanes in a fixed order.
constexpr int kAdjTG = 16;
__global__ __launch_bounds__(256) void k_real_adj_slp(RealAdjSameArgs a)
{
    constexpr int RN = SrcRec<0>::n;
    constexpr int TG = kAdjTG;
    // accumulators: target grad (2), source grad (2), density (2)
    constexpr int NV = 6;
    __shared__ double e1tab[SE2P_E1_NPIECE * SE2P_E1_ROW];
    __shared__ double etab[SE2P_EXP2_N];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < SE2P_E1_NPIECE * SE2P_E1_ROW; i += 256) e1tab[i] = c_e1_table[i];
    for (int i = threadIdx.x; i < SE2P_EXP2_N; i += blockDim.x) etab[i] = c_exp2_table[i];
    __syncthreads();

    const int b = xcd_chunk(a.chunk_off[a.ncell]);
    if (b < 0) return;   // uniform over the workgroup
    const int cell = a.block_cell[b];
    const int sub = 1 << (2 * kSubLog);
    const int l0 = a.off[cell * sub] + 64 * (b - a.chunk_off[cell]);
    const int l1 = a.off[(cell + 1) * sub];
    if (l0 + TG * wave >= l1) return;   // no target in this wavefront
    const int l = l0 + TG * wave + (lane % TG);
    const int row = lane / TG;   // source phase of the lane
    const bool active = l < l1;
    // lane point: position, densities, du
    double lx, ly, lq0 = 0.0, lq1 = 0.0, la1 = 0.0, la2 = 0.0;
    if (active) {
        const double *rp = a.rec + (int64_t)l * RN;
        const double4 v0 = *(const double4 *)rp;
        lx = v0.x;
        ly = v0.y;
        lq0 = v0.z;
        lq1 = v0.w;
        const double2 d = a.du[l];
        la1 = d.x;
        la2 = d.y;
    } else {
        lx = ly = __builtin_nan("");   // inactive lanes compute garbage, never stored
    }
    const int cx = cell % a.nsx, cy = cell / a.nsx;
    const double xi2 = a.xi2, cut2 = a.cut2;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;

    struct U {
        double x, y, q0, q1, b1, b2;
    };
    auto load = [&](int u) {
        U o;
        const double *rp = a.rec + (int64_t)u * RN;
        const double4 v0 = *(const double4 *)rp;
        o.x = v0.x;
        o.y = v0.y;
        o.q0 = v0.z;
        o.q1 = v0.w;
        const double2 d = a.du[u];
        o.b1 = d.x;
        o.b2 = d.y;
        return o;
    };
    // own: the own cell (no cutoff); img: a neighbour seen through a
    // periodic image (zx, zy not both zero; without one the differences are
    // the plain ones, up to the sign of a zero)
    auto pair = [&](const U &u, double zx, double zy, bool own, bool img) {
        double t1, t2, s1, s2;
        if (img) {
            diff<0>(u.x, u.y, lx, ly, zx, zy, t1, t2);   // lane as target
            s1 = lx - (u.x + zx);                         // lane as source
            s2 = ly - (u.y + zy);
        } else {
            t1 = u.x - lx;
            t2 = u.y - ly;
            s1 = lx - u.x;
            s2 = ly - u.y;
        }
        const double rt = se2p_sqsum(t1, t2), rs = se2p_sqsum(s1, s2);
        const bool in_t = own ? rt != 0.0 : rt < cut2;
        const bool in_s = own ? rs != 0.0 : rs < cut2;
        if (own && rs == 0.0) {   // coincident: self term, no gradient
            acc[4] += a.self * u.b1;
            acc[5] += a.self * u.b2;
        }
        if (!(in_t || in_s)) return;
        const double r2 = in_t ? rt : rs;
        const double x = r2 * xi2;
        const double e = se2p_exp_neg(x, etab);
        const double ir2 = rcp_nr(r2);
        if (in_t) slp_grad(t1, t2, xi2, e, ir2, la1, la2, u.q0, u.q1, acc[0], acc[1]);
        if (in_s) {
            slp_grad(s1, s2, xi2, e, ir2, u.b1, u.b2, lq0, lq1, acc[2], acc[3]);
            const double E = e1_lds(x, e, e1tab);
            const double hE = 0.5 * E, g = e * ir2;
            const double c = s1 * u.b2 - s2 * u.b1;
            acc[4] += hE * u.b1 + g * s2 * c;
            acc[5] += hE * u.b2 - g * s1 * c;
        }
    };
    // points ub + row, ub = u0, u0 + 4, ... < u1 (lanes past u1 idle; their
    // loads are clamped to u1 - 1)
    auto run_rows = [&](int u0, int u1, double zx, double zy, bool own, bool img) {
        if (u0 >= u1) return;
        int ub = u0;
        U cur = load(min(ub + row, u1 - 1));
        for (;;) {
            const int un = ub + 4;
            const U nxt = load(min(un + row, u1 - 1));
            if (ub + row < u1) pair(cur, zx, zy, own, img);
            if (un >= u1) break;
            cur = nxt;
            ub = un;
        }
    };
    run_rows(a.off[cell * sub], a.off[(cell + 1) * sub], 0.0, 0.0, true, false);
    double bx0, by0, bx1, by1;
    wave_bbox(active, lx, ly, bx0, by0, bx1, by1);
    const double lim = cut2 * (1.0 + 1e-9);
    const unsigned long long keep0 =
        keep_mask(a.box, 0, lane, 0, cx, cy, a.nsx, a.nsy, a.Lx, a.Ly, lim, bx0, by0, bx1, by1);
    const unsigned long long keep1 =
        keep_mask(a.box, 1, lane, 0, cx, cy, a.nsx, a.nsy, a.Lx, a.Ly, lim, bx0, by0, bx1, by1);
#pragma unroll 1
    for (int j = 0; j < 8; ++j) {
        unsigned bits = (unsigned)(((j < 4) ? keep0 : keep1) >> (16 * (j & 3))) & 0xffffu;
        if (bits == 0) continue;
        const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
        const int wx = (qx + a.nsx) % a.nsx, wy = (qy + a.nsy) % a.nsy;
        const int ub = wy * a.nsx + wx;
        const double zx = (a.Lx * (qx - wx)) / a.nsx;
        const double zy = (a.Ly * (qy - wy)) / a.nsy;
        while (bits) {
            const int ra = __builtin_ctz(bits);
            const int rb = ra + __builtin_ctz(~(bits >> ra));
            bits &= ~((1u << rb) - 1u);
            const int u0 = a.off[ub * sub + ra], u1 = a.off[ub * sub + rb];
            if (zx != 0.0 || zy != 0.0) run_rows(u0, u1, zx, zy, false, true);
            else run_rows(u0, u1, 0.0, 0.0, false, false);
        }
    }
    // rows 0..3: (row 0 + row 2) + (row 1 + row 3), in lanes 0..15
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        acc[v] += __shfl_down(acc[v], 32);
        acc[v] += __shfl_down(acc[v], 16);
    }
    if (row != 0 || !active) return;
    const int o = a.perm[l];
    const double s4 = 1.0 / (4 * SE2P_PI);
    if (a.dptar) {
        a.dptar[2 * o] = -acc[0] * s4;
        a.dptar[2 * o + 1] = -acc[1] * s4;
    }
    if (a.dpsrc) {
        a.dpsrc[2 * o] = acc[2] * s4;
        a.dpsrc[2 * o + 1] = acc[3] * s4;
    }
    if (a.df) {
        a.df[2 * o] = acc[4] * s4;
        a.df[2 * o + 1] = acc[5] * s4;
    }
}

// ---- windowed stresslet adjoint (targets are the sources) ------------------
// The stresslet adjoint pass over work items: a chunk with at least
// the scan's threshold of windows (k_real_win_scan, its own slot budget) is
// split into one item per window of its source sequence, whose 13 sums per
// lane go to slots that k_real_adj_win_final adds in window order; other
// chunks walk their whole sequence in one item.  Kept and rebuilt backward
// calls plan the same items from the same data, so their results agree bit
// for bit.
constexpr int kAdjNV = 13;          // stresslet adjoint sums per lane
constexpr int kAdjSlots = 16384;    // window slots (kAdjNV x 64 doubles each, 104 MiB)
constexpr int kAdjBlocksPerCU = 4;
constexpr int kAdjBlocksPerCUGroup = 3;   // when shares_gpu (as kWinBlocksPerCUGroup)

// The lane point's gradients from its 13 sums (stresslet): the lane's du
// and densities multiplied back in.
__device__ __forceinline__ void adj_dlp_out(const RealAdjSameArgs &a, int l, const double *acc)
{
    const double *rp = a.rec + (int64_t)l * SrcRec<1>::n;
    const double lq0 = -0.25 * rp[2], lq1 = -0.25 * rp[3], lq2 = -0.25 * rp[4];
    const double2 du = a.du[l];
    const double la1 = du.x, la2 = du.y;
    const int o = a.perm[l];
    const double s4 = 1.0 / (4 * SE2P_PI);
    if (a.dptar) {
        const double t0 = -0.25 * acc[0], t1 = -0.25 * acc[1], t2 = -0.25 * acc[2],
                     t3 = -0.25 * acc[3];
        a.dptar[2 * o] = -(la1 * t0 + la2 * t1) * s4;
        a.dptar[2 * o + 1] = -(la1 * t2 + la2 * t3) * s4;
    }
    if (a.dpsrc) {
        a.dpsrc[2 * o] = (lq0 * acc[4] + lq1 * acc[5] + lq2 * acc[6]) * s4;
        a.dpsrc[2 * o + 1] = (lq0 * acc[7] + lq1 * acc[8] + lq2 * acc[9]) * s4;
    }
    const double m1 = acc[10] * s4, m2 = acc[11] * s4, m3 = acc[12] * s4;
    const double f1 = a.f[2 * o], f2 = a.f[2 * o + 1];
    const double n1 = a.nv[2 * o], n2 = a.nv[2 * o + 1];
    if (a.df) {
        a.df[2 * o] = m1 * n1 + m2 * n2;
        a.df[2 * o + 1] = m2 * n1 + m3 * n2;
    }
    if (a.dn) {
        a.dn[2 * o] = m1 * f1 + m2 * f2;
        a.dn[2 * o + 1] = m2 * f1 + m3 * f2;
    }
}

__global__ __launch_bounds__(256, 4) void k_real_adj_win(RealAdjSameArgs a, WinPlan w)
{
    constexpr int RN = SrcRec<1>::n;
    constexpr int NV = kAdjNV;
    __shared__ double etab[SE2P_EXP2_N];
    __shared__ double red[kWaves - 1][NV][64];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < SE2P_EXP2_N; i += blockDim.x) etab[i] = c_exp2_table[i];
    const int total = w.meta[0];
    const int per = (total + 7) >> 3;
    const int xcd = blockIdx.x & 7;
    const int sub = 1 << (2 * kSubLog);
    const double xi2 = a.xi2, cut2 = a.cut2;
    int qi = 0;   // thread 0: queues found empty so far
    for (;;) {
        if (threadIdx.x == 0) {
            int item = -1;
            for (; qi < 8; ++qi) {
                const int x = (xcd + qi) & 7;
                const int lo = x * per, nq = min(per, total - lo);
                if (nq <= 0) continue;
                const int it = atomicAdd(&w.meta[kQueueStride * (x + 1)], 1);
                if (it < nq) {
                    item = lo + it;
                    break;
                }
            }
            s_item = item;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item < 0) break;   // uniform over the workgroup

        const int c = w.item_chunk[item];
        const int sl = w.slot[c];
        const int cell = a.block_cell[c];
        const int l = a.off[cell * sub] + 64 * (c - a.chunk_off[cell]) + lane;
        const bool active = l < a.off[(cell + 1) * sub];
        double lx, ly;
        if (active) {
            const double2 pl = *(const double2 *)(a.rec + (int64_t)l * RN);
            lx = pl.x;
            ly = pl.y;
        } else {
            lx = ly = __builtin_nan("");   // inactive lanes compute garbage, never stored
        }
        const int cx = cell % a.nsx, cy = cell / a.nsx;
        const int n = w.npos[c];
        const int kb = sl >= 0 ? item - w.item_off[c] : 0;
        const int wa = sl >= 0 ? kb * kWin : 0, wb = sl >= 0 ? min(wa + kWin, n) : n;
        const unsigned long long keep0 = w.keep[2 * c], keep1 = w.keep[2 * c + 1];
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;

        struct U {
            double x, y, q0, q1, q2, b1, b2;
        };
        auto load = [&](int u) {
            U o;
            const double *rp = a.rec + (int64_t)u * RN;
            const double4 v0 = *(const double4 *)rp;
            o.x = v0.x;
            o.y = v0.y;
            o.q0 = v0.z;
            o.q1 = v0.w;
            o.q2 = rp[4];
            const double2 d = a.du[u];
            o.b1 = d.x;
            o.b2 = d.y;
            return o;
        };
        // the stresslet pair (see above); img: a neighbour seen through a
        // periodic image (without one the difference is the plain one, up to
        // the sign of a zero)
        auto pair = [&](const U &u, double zx, double zy, bool own, bool img) {
            double t1, t2;
            if (img) {
                diff<1>(u.x, u.y, lx, ly, zx, zy, t1, t2);
            } else {
                t1 = u.x - lx;
                t2 = u.y - ly;
            }
            double d11, d22;
            const double r2 = se2p_sqsum2(t1, t2, &d11, &d22);
            if (own ? (r2 == 0.0) : !(r2 < cut2)) return;
            const double x = xi2 * r2;
            const double e2 = se2p_exp_neg(x, etab);
            const double ir2 = rcp_nr(r2);
            const double ir4 = ir2 * ir2;
            const double al = -4.0 * e2 * (1.0 + x) * ir4;
            const double be = 2.0 * xi2 * e2;
            const double g1 = 8.0 * e2 * (x * (x + 2.0) + 2.0) * (ir4 * ir2);
            const double g2 = -2.0 * xi2 * be;
            const double A11 = al * d11, A22 = al * d22, A12 = al * (t1 * t2);
            const double b1 = A11 + be, b3 = A22 + be;
            const double tt0 = t1 * (b1 + 2.0 * be), tt1 = t2 * b1, tt2 = t1 * b3,
                         tt3 = t2 * (b3 + 2.0 * be);
            const double G11 = g1 * d11, G22 = g1 * d22;
            const double h0 = t1 * (G11 + 3.0 * g2), h1 = t2 * (G11 + g2), h2 = t1 * (G22 + g2),
                         h3 = t2 * (G22 + 3.0 * g2);
            const double D10 = t1 * h0 + 3.0 * b1, D11 = t1 * h1 + 2.0 * A12, D12 = t1 * h2 + b3,
                         D13 = t1 * h3;
            const double D20 = t2 * h0, D21 = t2 * h1 + b1, D22 = t2 * h2 + 2.0 * A12,
                         D23 = t2 * h3 + 3.0 * b3;
            acc[0] = fma(D12, u.q2, fma(D11, u.q1, fma(D10, u.q0, acc[0])));
            acc[1] = fma(D13, u.q2, fma(D12, u.q1, fma(D11, u.q0, acc[1])));
            acc[2] = fma(D22, u.q2, fma(D21, u.q1, fma(D20, u.q0, acc[2])));
            acc[3] = fma(D23, u.q2, fma(D22, u.q1, fma(D21, u.q0, acc[3])));
            acc[4] = fma(D11, u.b2, fma(D10, u.b1, acc[4]));
            acc[5] = fma(D12, u.b2, fma(D11, u.b1, acc[5]));
            acc[6] = fma(D13, u.b2, fma(D12, u.b1, acc[6]));
            acc[7] = fma(D21, u.b2, fma(D20, u.b1, acc[7]));
            acc[8] = fma(D22, u.b2, fma(D21, u.b1, acc[8]));
            acc[9] = fma(D23, u.b2, fma(D22, u.b1, acc[9]));
            acc[10] = fma(-u.b2, tt1, fma(-u.b1, tt0, acc[10]));
            acc[11] = fma(-u.b2, tt2, fma(-u.b1, tt1, acc[11]));
            acc[12] = fma(-u.b2, tt3, fma(-u.b1, tt2, acc[12]));
        };
        auto run = [&](int u, int u1, double zx, double zy, bool own, bool img) {
            for (; u + kWaves < u1; u += 2 * kWaves) {
                const U c0 = load(u), c1 = load(u + kWaves);
                pair(c0, zx, zy, own, img);
                pair(c1, zx, zy, own, img);
            }
            if (u < u1) pair(load(u), zx, zy, own, img);
        };
        // the walk over [wa, wb) of the sequence (k_real_win's walk)
        int j = -1, s0 = a.off[cell * sub], s1 = a.off[(cell + 1) * sub], pos = 0, sb = 0;
        unsigned bits = 0;
        double zx = 0.0, zy = 0.0;
        bool more = true;
        while (more) {
            const int pe = pos + (s1 - s0);
            const int pa = max(pos, wa), pb = min(pe, wb);
            if (pa < pb) {
                const int sa = s0 + (pa - pos) + ((wave - pa) & (kWaves - 1));
                const int se = s0 + (pb - pos);
                if (j < 0) run(sa, se, 0.0, 0.0, true, false);
                else if (zx == 0.0 && zy == 0.0) run(sa, se, 0.0, 0.0, false, false);
                else run(sa, se, zx, zy, false, true);
            }
            if (pe >= wb) break;
            pos = pe;
            while (bits == 0) {
                if (++j == 8) {
                    more = false;
                    break;
                }
                bits = (unsigned)(((j < 4) ? keep0 : keep1) >> (16 * (j & 3))) & 0xffffu;
                if (bits) {
                    const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
                    const int wx = (qx + a.nsx) % a.nsx, wy = (qy + a.nsy) % a.nsy;
                    sb = wy * a.nsx + wx;
                    zx = (a.Lx * (qx - wx)) / a.nsx;   // :248-249
                    zy = (a.Ly * (qy - wy)) / a.nsy;
                }
            }
            if (!more) break;
            const int ra = __builtin_ctz(bits);
            const int rb = ra + __builtin_ctz(~(bits >> ra));
            bits &= ~((1u << rb) - 1u);
            s0 = a.off[sb * sub + ra];
            s1 = a.off[sb * sub + rb];
        }
        // the four wavefronts' sums in wave order
        if (wave > 0)
#pragma unroll
            for (int v = 0; v < NV; ++v) red[wave - 1][v][lane] = acc[v];
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int q = 0; q < kWaves - 1; ++q) acc[v] += red[q][v][lane];
            if (sl >= 0) {
                double *pp = w.part + (int64_t)(sl + kb) * (NV * 64) + lane;
#pragma unroll
                for (int v = 0; v < NV; ++v) pp[v * 64] = acc[v];
            } else if (active) {
                adj_dlp_out(a, l, acc);
            }
        }
        __syncthreads();   // red and s_item are reused
    }
}

// Split chunks of k_real_adj_win: window sums added in window order from
// zero, then the lane's gradients.  One wavefront per split chunk.
__global__ __launch_bounds__(256) void k_real_adj_win_final(RealAdjSameArgs a, WinPlan w)
{
    constexpr int NV = kAdjNV;
    const int lane = threadIdx.x & 63;
    const int nsplit = w.meta[2];
    const int sub = 1 << (2 * kSubLog);
    for (int i = blockIdx.x * kWaves + (threadIdx.x >> 6); i < nsplit; i += gridDim.x * kWaves) {
        const int4 r = w.split[i];   // (slot, windows, first lane point, cell)
        const int l = r.z + lane;
        if (l >= a.off[(r.w + 1) * sub]) continue;
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        const double *pp = w.part + (int64_t)r.x * (NV * 64) + lane;
        for (int k = 0; k < r.y; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += pp[(int64_t)k * (NV * 64) + v * 64];
        adj_dlp_out(a, l, acc);
    }
}

// du in sorted order, one double2 per point.
__global__ void k_gather_du(const double *__restrict__ v, const int *__restrict__ perm, int n,
                            double2 *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    out[i] = make_double2(v[2 * o], v[2 * o + 1]);
}

__global__ void k_gather_pairs(const double *__restrict__ v, const int *__restrict__ perm, int n,
                               double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    out[i] = v[2 * o];
    out[n + i] = v[2 * o + 1];
}

// ---- box-table gathers -----------------------------------------------------
// Global point i belongs to box pt_box[i] (all to box 0 when pt_box is null).
__device__ __forceinline__ int box_of(const int *pt_box, int i) { return pt_box ? pt_box[i] : 0; }

// (global cell, Morton sub-cell) key of source (which = 0) or target (1) i.
__device__ __forceinline__ int box_cell_key(const BoxTable &boxes, const int *pt_box, int which,
                                            int i)
{
    const RealBox &B = boxes.get(box_of(pt_box, i));
    const int l = i - (which ? B.t_base : B.s_base);
    const double *p = which ? B.ptar : B.psrc;
    return (B.cell_base << (2 * kSubLog)) +
           se2p_cell_key(p[2 * l], p[2 * l + 1], B.Lx, B.Ly, B.nsx, B.nsy, kSubLog);
}

__global__ void k_box_cell_keys(const BoxTable boxes, const int *__restrict__ pt_box,
                                int n, int which, int *__restrict__ key)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = box_cell_key(boxes, pt_box, which, i);
}

// sources: position of input point g at sorted slot i and, with one
// right-hand side (rec != null), its record (SrcRec).  Several right-hand
// sides: the densities are written by k_gather_densities.
template <int KIND>
__device__ __forceinline__ void gather_source(const BoxTable &boxes, const int *pt_box, int g, int i,
                                              double2 *pos, double *rec)
{
    const RealBox &B = boxes.get(box_of(pt_box, g));
    const int o = g - B.s_base;
    const double x = B.psrc[2 * o], y = B.psrc[2 * o + 1];
    pos[i] = make_double2(x, y);
    if (rec) {
        double *rp = rec + (int64_t)i * SrcRec<KIND>::n;
        const double f1 = B.f[2 * o], f2 = B.f[2 * o + 1];
        if (KIND == 0) {
            *(double4 *)rp = make_double4(x, y, f1, f2);
        } else {
            const double n1 = B.nv[2 * o], n2 = B.nv[2 * o + 1];
            const double s0 = f1 * n1, s1 = f1 * n2 + f2 * n1, s2 = f2 * n2;
            // the pair kernel's constant factors folded in: -4 S (exactly
            // -4 times, a power of two) and 2 xi^2 times (c1, c2, s1)
            const double b = 2.0 * B.xi2;
            *(double4 *)rp = make_double4(x, y, -4.0 * s0, -4.0 * s1);
            *(double4 *)(rp + 4) =
                make_double4(-4.0 * s2, b * (3.0 * s0 + s2), b * (s0 + 3.0 * s2), b * s1);
        }
        return;
    }
}

// Several right-hand sides: the densities of sorted slot i and right-hand
// side r, source-major, q[(i * nrhs + r) * ncomp + c], so that a source's
// values for a group of right-hand sides are contiguous.  One thread per
// (i, r), consecutive threads writing consecutive rows.
template <int KIND>
__global__ void k_gather_densities(const BoxTable boxes, const int *__restrict__ pt_box,
                                   const int *__restrict__ perm, int n, int nrhs, int64_t dstride,
                                   double *__restrict__ q)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nrhs) return;
    const int i = (int)(t / nrhs), r = (int)(t % nrhs);
    const int g = perm[i];
    const RealBox &B = boxes.get(box_of(pt_box, g));
    const int o = g - B.s_base;
    const double *fr = B.f + (int64_t)r * dstride;
    const double f1 = fr[2 * o], f2 = fr[2 * o + 1];
    if (KIND == 0) {
        q[2 * t] = f1;
        q[2 * t + 1] = f2;
    } else {
        const double n1 = B.nv[2 * o], n2 = B.nv[2 * o + 1];
        // S = f (x) n (mex_stokes_dlp_real.cpp:86-89); S12 + S21 merged,
        // then the two combinations the pair kernel uses
        const double s0 = f1 * n1, s1 = f1 * n2 + f2 * n1, s2 = f2 * n2;
        double *qr = q + kDlpComp * t;
        qr[0] = s0;
        qr[1] = s1;
        qr[2] = s2;
        qr[3] = 3.0 * s0 + s2;
        qr[4] = s0 + 3.0 * s2;
    }
}

template <int KIND>
__global__ void k_box_gather_sources(const BoxTable boxes,
                                     const int *__restrict__ pt_box, const int *__restrict__ perm,
                                     int n, double2 *__restrict__ pos, double *__restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gather_source<KIND>(boxes, pt_box, perm[i], i, pos, rec);
}

__device__ __forceinline__ void gather_target(const BoxTable &boxes, const int *pt_box, int g, int i,
                                              double2 *out)
{
    const RealBox &B = boxes.get(box_of(pt_box, g));
    const int o = g - B.t_base;
    out[i] = make_double2(B.ptar[2 * o], B.ptar[2 * o + 1]);
}

__global__ void k_box_gather_targets(const BoxTable boxes,
                                     const int *__restrict__ pt_box, const int *__restrict__ perm,
                                     int n, double2 *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gather_target(boxes, pt_box, perm[i], i, out);
}

// Fused-sort functors (se2p_fsort.h): cell keys, and the gathers run by the
// final scatter.
struct CellKeyFn {
    BoxTable boxes;
    const int *pt_box;
    int which;
    __device__ int operator()(int i) const { return box_cell_key(boxes, pt_box, which, i); }
};

// Sources: also the bounding box of every key run (as k_run_bbox).
template <int KIND>
struct SrcGather {
    static constexpr bool kBox = true;
    BoxTable boxes;
    const int *pt_box;
    double2 *pos;
    double *rec;
    double4 *box;
    __device__ double2 operator()(int i, int g, int) const
    {
        gather_source<KIND>(boxes, pt_box, g, i, pos, rec);
        const RealBox &B = boxes.get(box_of(pt_box, g));
        const int o = g - B.s_base;
        return make_double2(B.psrc[2 * o], B.psrc[2 * o + 1]);
    }
};

struct TgtGather {
    static constexpr bool kBox = false;
    BoxTable boxes;
    const int *pt_box;
    double2 *pos;
    __device__ void operator()(int i, int g, int) const { gather_target(boxes, pt_box, g, i, pos); }
};


// ---- gathers ------------------------------------------------------------
__global__ void k_gather_targets(const double *__restrict__ p, const int *__restrict__ perm, int n,
                                 double2 *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    out[i] = make_double2(p[2 * o], p[2 * o + 1]);
}

// sources: positions and per-RHS densities in cell order
template <int KIND>
__global__ void k_gather_sources(const double *__restrict__ p, const double *__restrict__ f,
                                 const double *__restrict__ nv, const int *__restrict__ perm,
                                 int n, int nrhs, int64_t dstride, double2 *__restrict__ pos,
                                 double *__restrict__ q)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    pos[i] = make_double2(p[2 * o], p[2 * o + 1]);
    if (KIND == 0) {
        for (int r = 0; r < nrhs; ++r) {
            const double *fr = f + (int64_t)r * dstride;
            q[(int64_t)(2 * r) * n + i] = fr[2 * o];
            q[(int64_t)(2 * r + 1) * n + i] = fr[2 * o + 1];
        }
    } else {
        double n1 = nv[2 * o], n2 = nv[2 * o + 1];
        for (int r = 0; r < nrhs; ++r) {
            const double *fr = f + (int64_t)r * dstride;
            double f1 = fr[2 * o], f2 = fr[2 * o + 1];
            // S = f (x) n (mex_stokes_dlp_real.cpp:86-89); S12 + S21 merged
            q[(int64_t)(3 * r) * n + i] = f1 * n1;
            q[(int64_t)(3 * r + 1) * n + i] = f1 * n2 + f2 * n1;
            q[(int64_t)(3 * r + 2) * n + i] = f2 * n2;
        }
    }
}

static int upload_table(se2p_ctx *ctx)
{
    static std::mutex mu;
    static bool done[64] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (ctx->device >= 0 && ctx->device < 64 && done[ctx->device]) return 0;
    static double host[SE2P_E1_NPIECE * SE2P_E1_ROW];
    for (int p = 0; p < SE2P_E1_NPIECE; ++p) {
        double *row = host + p * SE2P_E1_ROW;
        row[0] = se2p_e1_scale_shift[p][0];
        row[1] = se2p_e1_scale_shift[p][1];
        for (int k = 0; k < SE2P_E1_NCOEF; ++k) row[2 + k] = se2p_e1_coef[p][k];
        for (int k = 2 + SE2P_E1_NCOEF; k < SE2P_E1_ROW; ++k) row[k] = 0.0;
    }
    SE2P_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_e1_table), host, sizeof(host)));
    if (ctx->device >= 0 && ctx->device < 64) done[ctx->device] = true;
    return 0;
}

// ---- host driver ----------------------------------------------------------
// Sorted state of one real-space call: the cell lists of the sources (and of
// the targets unless they are the sources), the gathered source data, the
// chunk plans of the targets and the run boxes of the sources.
// real_layout carves it, real_prepare fills it (the sort stage), the pair
// kernels read it.
struct RealPlan {
    int kind = 0, nbox = 0, nrhs = 1;
    int64_t ns = 0, nt = 0, ncell = 0, dstride = 0;
    int nkeys = 0;
    bool same_points = false;
    int64_t max_chunks = 0;
    unsigned grid8 = 0;
    KeySort kss{}, kst{};
    double2 *spos = nullptr, *tpos = nullptr;
    double *q = nullptr, *rec = nullptr;
    int *coff = nullptr, *bmap = nullptr, *offs = nullptr, *s_box = nullptr, *t_box = nullptr,
        *cell_box = nullptr;
    unsigned long long *cnt = nullptr;
    double4 *sbox = nullptr;
    RealBox *dbox = nullptr;
    BoxTable btab{};
    WinPlan win{};   // single density: windows, work items and split slots
    int win_slots = 0;
};

// Box bases, derived parameters and the plan's sizes; 0 or an error code.
static int real_plan(se2p_ctx *ctx, int kind, RealBox *boxes, int nbox, int nrhs, int64_t dstride,
                     RealPlan *p)
{
    const int sub = 1 << (2 * kSubLog);
    int64_t ns = 0, nt = 0, ncell = 0;
    for (int b = 0; b < nbox; ++b) {
        RealBox &B = boxes[b];
        B.s_base = (int)ns;
        B.t_base = (int)nt;
        B.cell_base = (int)ncell;
        B.xi2 = B.xi * B.xi;
        B.cut2 = B.Lx * B.Ly / B.nsx / B.nsy;                 // mex_stokes_slp_real.cpp:127
        B.self = -(0.5 * SE2P_EULER_GAMMA + 1.0) - log(B.xi);  // :124
        ns += B.ns;
        nt += B.nt;
        ncell += (int64_t)B.nsx * B.nsy;
    }
    if (ns > 0x7fffffffLL || nt > 0x7fffffffLL || ncell * sub > 0x7fffffffLL)
        return fail(ctx, SE2P_ERR_ARG, "too many points or cells in one call");
    *p = RealPlan{};
    p->kind = kind;
    p->nbox = nbox;
    p->nrhs = nrhs;
    p->dstride = dstride;
    p->ns = ns;
    p->nt = nt;
    p->ncell = ncell;
    // targets that are the source array itself share the sources' cell lists
    p->same_points = nbox == 1 && boxes[0].ptar == boxes[0].psrc && boxes[0].nt == boxes[0].ns;
    p->nkeys = (int)(ncell * sub);
    // upper bound on chunks (every non-empty cell wastes < 1 chunk); the
    // cell-list kernels want a multiple of 8 workgroups (XCD order)
    p->max_chunks = (nt + 63) / 64 + ncell;
    p->grid8 = (unsigned)((p->max_chunks + 7) & ~int64_t(7));
    return 0;
}

static void real_layout(Carve &cv, RealPlan &p)
{
    const int ncomp = p.kind == 0 ? 2 : kDlpComp;
    char *c = cv.take<char>(keysort_bytes(p.ns, p.nkeys));
    if (c) keysort_carve(c, p.ns, p.nkeys, &p.kss);
    if (!p.same_points) {
        c = cv.take<char>(keysort_bytes(p.nt, p.nkeys));
        if (c) keysort_carve(c, p.nt, p.nkeys, &p.kst);
    }
    p.spos = cv.take<double2>(p.ns);
    p.tpos = p.same_points ? p.spos : cv.take<double2>(p.nt);
    if (p.nrhs == 1) p.rec = cv.take<double>((size_t)(p.kind == 0 ? SrcRec<0>::n : SrcRec<1>::n) * p.ns);
    else p.q = cv.take<double>((size_t)ncomp * p.nrhs * p.ns);
    p.coff = cv.take<int>(p.ncell + 1);
    p.cnt = cv.take<unsigned long long>(2);
    p.bmap = cv.take<int>(p.max_chunks);
    p.sbox = cv.take<double4>(p.nkeys);
    if (p.nrhs == 1) {
        p.win.slots = p.win_slots;
        p.win.keep = cv.take<unsigned long long>(2 * (size_t)p.max_chunks);
        p.win.npos = cv.take<int>(p.max_chunks);
        p.win.where = cv.take<int2>(p.max_chunks);
        p.win.item_off = cv.take<int>(p.max_chunks);
        p.win.slot = cv.take<int>(p.max_chunks);
        p.win.item_chunk = cv.take<int>(p.max_chunks + p.win_slots);
        p.win.split = cv.take<int4>(p.max_chunks);
        p.win.meta = cv.take<int>(kQueueStride * 9);
        p.win.part = cv.take<double>(128 * (size_t)p.win_slots);
    }
    if (p.nbox > 1) {
        p.dbox = cv.take<RealBox>(p.nbox);
        p.offs = cv.take<int>(3 * (size_t)(p.nbox + 1));
        p.s_box = cv.take<int>(p.ns);
        p.t_box = cv.take<int>(p.nt);
        p.cell_box = cv.take<int>(p.ncell);
    }
}

// The sort stage: box tables, cell lists, gathers, chunk plans, run boxes.
static int real_prepare(se2p_ctx *ctx, const RealBox *boxes, RealPlan &p)
{
    hipStream_t st = ctx->stream;
    const int sub = 1 << (2 * kSubLog);
    const int threads = 256;
    const int64_t ns = p.ns, nt = p.nt, ncell = p.ncell;
    const int nkeys = p.nkeys, nbox = p.nbox, nrhs = p.nrhs, kind = p.kind;
    const int64_t dstride = p.dstride;
    if (nbox == 1) {
        p.btab.tab = nullptr;
        p.btab.one = boxes[0];
    } else {
        p.btab.tab = p.dbox;
        // box table and the source / target / cell offsets of every box,
        // written by kernels that carry the records as arguments: no host
        // copy enters the stream, so a captured sequence replays on its own
        for (int b0 = 0; b0 < nbox; b0 += kBoxesPerLaunch) {
            BoxBatch bb;
            bb.first = b0;
            bb.count = std::min(kBoxesPerLaunch, nbox - b0);
            for (int i = 0; i < bb.count; ++i) bb.b[i] = boxes[b0 + i];
            hipLaunchKernelGGL(k_write_boxes, dim3(1), dim3(kBoxesPerLaunch), 0, st, bb, p.dbox,
                               p.offs, nbox, (int)ns, (int)nt, (int)ncell);
        }
        expand_map(st, p.offs, nbox, (int)ns, p.s_box);
        expand_map(st, p.offs + nbox + 1, nbox, (int)nt, p.t_box);
        expand_map(st, p.offs + 2 * (nbox + 1), nbox, (int)ncell, p.cell_box);
    }
    const BoxTable &btab = p.btab;
    // Sources and targets in (cell, Morton sub-cell) order, stable: a cell's
    // sources are summed in a fixed order on every run.  The fused sort
    // (se2p_fsort.h) also gathers the sorted data and plans the chunks of 64
    // targets; otherwise separate launches do.
    const KeySort &tks = p.same_points ? p.kss : p.kst;
    const PlanArgs plan{nullptr, tks.off, sub, 64, (int)ncell, p.coff, p.bmap};
    bool planned = false, boxed = false;
    const bool box_kernel = (ctx->sort_mode & SE2P_SORT_BOX_KERNEL) != 0;
    if (fsort_ok(ctx, ns, nkeys)) {
        FsPlanExtra x0{plan, plan, p.same_points ? 1 : 0};
        const CellKeyFn kf{btab, p.s_box, 0};
        int rc;
        if (kind == 0)
            rc = fsort_run(ctx, &p.kss, ns, nkeys, kf,
                           SrcGather<0>{btab, p.s_box, p.spos, p.rec, p.sbox}, x0, FsNoExtra{});
        else
            rc = fsort_run(ctx, &p.kss, ns, nkeys, kf,
                           SrcGather<1>{btab, p.s_box, p.spos, p.rec, p.sbox}, x0, FsNoExtra{});
        if (rc) return rc;
        planned = p.same_points;
        boxed = !box_kernel;
    } else {
        if (ns > 0)
            hipLaunchKernelGGL(k_box_cell_keys, dim3((unsigned)((ns + threads - 1) / threads)),
                               dim3(threads), 0, st, btab, p.s_box, (int)ns, 0, p.kss.key);
        if (int rc = keysort_run(ctx, &p.kss, ns, nkeys)) return rc;
        if (ns > 0) {
            const unsigned blocks = (unsigned)((ns + threads - 1) / threads);
            if (kind == 0)
                hipLaunchKernelGGL(k_box_gather_sources<0>, dim3(blocks), dim3(threads), 0, st, btab,
                                   p.s_box, p.kss.perm, (int)ns, p.spos, p.rec);
            else
                hipLaunchKernelGGL(k_box_gather_sources<1>, dim3(blocks), dim3(threads), 0, st, btab,
                                   p.s_box, p.kss.perm, (int)ns, p.spos, p.rec);
        }
    }
    if (!p.same_points) {
        if (fsort_ok(ctx, nt, nkeys)) {
            if (int rc = fsort_run(ctx, &p.kst, nt, nkeys, CellKeyFn{btab, p.t_box, 1},
                                   TgtGather{btab, p.t_box, p.tpos}, FsPlanExtra{plan, plan, 1},
                                   FsNoExtra{}))
                return rc;
            planned = true;
        } else {
            hipLaunchKernelGGL(k_box_cell_keys, dim3((unsigned)((nt + threads - 1) / threads)),
                               dim3(threads), 0, st, btab, p.t_box, (int)nt, 1, p.kst.key);
            if (int rc = keysort_run(ctx, &p.kst, nt, nkeys)) return rc;
            hipLaunchKernelGGL(k_box_gather_targets, dim3((unsigned)((nt + threads - 1) / threads)),
                               dim3(threads), 0, st, btab, p.t_box, tks.perm, (int)nt, p.tpos);
        }
    }
    if (p.q && ns > 0) {   // several right-hand sides
        const unsigned blocks = (unsigned)(((int64_t)ns * nrhs + threads - 1) / threads);
        if (kind == 0)
            hipLaunchKernelGGL(k_gather_densities<0>, dim3(blocks), dim3(threads), 0, st, btab, p.s_box,
                               p.kss.perm, (int)ns, nrhs, dstride, p.q);
        else
            hipLaunchKernelGGL(k_gather_densities<1>, dim3(blocks), dim3(threads), 0, st, btab, p.s_box,
                               p.kss.perm, (int)ns, nrhs, dstride, p.q);
    }
    if (!planned) plan_chunks(st, nullptr, tks.off, sub, 64, (int)ncell, p.coff, p.bmap);   // chunks of 64 targets
    if (!boxed)
        hipLaunchKernelGGL(k_run_bbox, dim3((nkeys + 3) / 4), dim3(256), 0, st, p.spos, p.kss.off,
                           nkeys, p.sbox);
    SE2P_HIP(hipGetLastError());
    return 0;
}

// State a forward call keeps for its backward call (se2p_set_keep): one
// single-box, single-density, targets-are-sources call, identified by its
// arguments.  The plan's arrays live in a buffer of their own.
struct KeptReal {
    char *buf = nullptr;
    size_t bytes = 0;
    bool valid = false;
    int kind = 0;
    const double *psrc = nullptr, *f = nullptr, *nv = nullptr;
    int64_t n = 0;
    double xi = 0, Lx = 0, Ly = 0;
    int nsx = 0, nsy = 0;
    RealPlan plan;
    bool match(int k, const RealGeom &g, const double *ff, const double *nn) const
    {
        return valid && k == kind && g.psrc == psrc && g.ptar == psrc && g.ns == n && g.nt == n &&
               ff == f && (k == 0 || nn == nv) && g.xi == xi && g.Lx == Lx && g.Ly == Ly &&
               g.nsx == nsx && g.nsy == nsy;
    }
};

void free_kept_real(se2p_ctx *ctx)
{
    if (!ctx->kept_real) return;
    if (ctx->kept_real->buf) (void)hipFree(ctx->kept_real->buf);
    delete ctx->kept_real;
    ctx->kept_real = nullptr;
}

// Real-space sum over nbox independent boxes in one pass: one sort of all
// points by (global cell, sub-cell), one gather, one launch of the pair
// kernel.  boxes[] are host records; their bases and derived parameters are
// filled in here.
int real_sum_boxes(se2p_ctx *ctx, int kind, RealBox *boxes, int nbox, int nrhs, int64_t dstride,
                   int64_t ustride, int64_t *pair_counts, char *fixed, size_t fixed_bytes)
{
    hipStream_t st = ctx->stream;
    RealPlan p;
    if (int rc = real_plan(ctx, kind, boxes, nbox, nrhs, dstride, &p)) return rc;
    const bool counting = pair_counts != nullptr;
    if (p.nt == 0) return 0;
    if (kind == 0 && upload_table(ctx)) return SE2P_ERR_HIP;
    // a call whose backward may follow keeps its plan in a buffer of its own
    const bool keep = ctx->keep && !ctx->host_ptrs && nbox == 1 && nrhs == 1 && !counting &&
                      p.same_points && !fixed;
    p.win_slots = nrhs == 1 && !counting ? kWinSlots : 0;
    Carve sizing;
    real_layout(sizing, p);
    if (fixed && sizing.off > fixed_bytes && p.win_slots > 0) {
        // the fixed scratch holds fewer window slots (results do not depend
        // on how many chunks are split)
        p.win_slots = 0;
        Carve base;
        real_layout(base, p);
        if (base.off <= fixed_bytes) {
            const size_t spare = fixed_bytes - base.off;
            const size_t per_slot = 128 * sizeof(double) + sizeof(int);   // sums + item entry
            p.win_slots = spare > 1024 ? (int)std::min<size_t>(kWinSlots, (spare - 1024) / per_slot) : 0;
        }
        sizing = Carve{};
        real_layout(sizing, p);
    }
    Carve cv;
    if (fixed) {
        // caller-provided scratch: nothing is allocated
        if (sizing.off > fixed_bytes)
            return fail(ctx, SE2P_ERR_ARG,
                        "the sorted state needs %zu bytes of scratch; the fixed scratch holds %zu",
                        sizing.off, fixed_bytes);
        cv.base = fixed;
    } else if (keep) {
        if (!ctx->kept_real) ctx->kept_real = new KeptReal;
        KeptReal &kr = *ctx->kept_real;
        kr.valid = false;
        if (sizing.off > kr.bytes) {
            if (int rc = refuse_if_frozen(ctx, "growing the kept real-space state", sizing.off))
                return rc;
            if (kr.buf) {
                SE2P_HIP(hipStreamSynchronize(st));
                SE2P_HIP(hipFree(kr.buf));
                kr.buf = nullptr;
                kr.bytes = 0;
            }
            SE2P_HIP(hipMalloc((void **)&kr.buf, sizing.off));
            kr.bytes = sizing.off;
        }
        cv.base = kr.buf;
    } else {
        if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
        cv.base = ctx->ws;
    }
    real_layout(cv, p);

    stage_begin(ctx, ST_SORT);
    if (int rc = real_prepare(ctx, boxes, p)) return rc;
    stage_end(ctx, ST_SORT);
    if (keep) {
        KeptReal &kr = *ctx->kept_real;
        const RealBox &B = boxes[0];
        kr.kind = kind;
        kr.psrc = B.psrc;
        kr.f = B.f;
        kr.nv = B.nv;
        kr.n = B.ns;
        kr.xi = B.xi;
        kr.Lx = B.Lx;
        kr.Ly = B.Ly;
        kr.nsx = B.nsx;
        kr.nsy = B.nsy;
        kr.plan = p;
        kr.valid = true;
    }

    const KeySort &tks = p.same_points ? p.kss : p.kst;
    RealArgs a{};
    a.tpos = p.tpos;
    a.tperm = tks.perm;
    a.toff = tks.off;
    a.chunk_off = p.coff;
    a.block_cell = p.bmap;
    a.spos = p.spos;
    a.q = p.q;
    a.soff = p.kss.off;
    a.ns = p.ns;
    a.ncell = (int)p.ncell;
    a.nrhs = nrhs;
    a.ustride = ustride;
    a.boxes = p.btab;
    a.cell_box = p.cell_box;
    a.pair_counts = p.cnt;
    const double2 *spos = p.spos;
    const double *q = p.q, *rec = p.rec;
    const int *soff = p.kss.off;
    const double4 *sbox = p.sbox;

    stage_begin(ctx, ST_REAL);
    if (counting) {
        SE2P_HIP(hipMemsetAsync(p.cnt, 0, 2 * sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_real_count, dim3((unsigned)p.max_chunks), dim3(64), 0, st, a);
    } else {
        // groups of kBatchRB, then of four right-hand sides share every pair
        // evaluation; the remainder goes one at a time
        const int n8 = kBatchRB == 8 ? nrhs / 8 * 8 : 0;
        if (n8 > 0) {
            a.r_begin = 0;
            a.r_end = n8;
            const dim3 grid((unsigned)(p.max_chunks * (n8 / 8)));
            if (kind == 0) hipLaunchKernelGGL((k_real<0, 8>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
            else hipLaunchKernelGGL((k_real<1, 8>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
        }
        const int n4 = n8 + (nrhs - n8) / 4 * 4;
        if (n4 > n8) {
            a.r_begin = n8;
            a.r_end = n4;
            const dim3 grid((unsigned)(p.max_chunks * ((n4 - n8) / 4)));
            if (kind == 0) hipLaunchKernelGGL((k_real<0, 4>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
            else hipLaunchKernelGGL((k_real<1, 4>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
        }
        if (n4 < nrhs) {
            a.r_begin = n4;
            a.r_end = nrhs;
            const dim3 grid((unsigned)(p.max_chunks * (nrhs - n4)));
            if (rec) {   // single density: source records, windows of the chunks
                const WinPlan &wp = p.win;
                const unsigned cgrid = (unsigned)((p.max_chunks + kWaves - 1) / kWaves);
                hipLaunchKernelGGL(k_real_win_plan, dim3(cgrid), dim3(256), 0, st, a, soff, sbox, wp);
                hipLaunchKernelGGL(k_real_win_scan, dim3(1), dim3(64 * kWinScanWaves), 0, st, p.coff,
                                   (int)p.ncell, wp);
                const int64_t items = p.max_chunks + p.win_slots;
                const int bpc = shares_gpu(ctx) ? kWinBlocksPerCUGroup : kWinBlocksPerCU;
                const int64_t cap = (int64_t)ctx->ncu * bpc;
                const unsigned wgrid = (unsigned)((std::min(items, cap) + 7) & ~int64_t(7));
                const unsigned fgrid = std::min<unsigned>(cgrid, kWinFinalBlocks);
                stage_begin(ctx, ST_PAIR);
                if (kind == 0) hipLaunchKernelGGL(k_real_win<0>, dim3(wgrid), dim3(256), 0, st, a, rec, soff, wp);
                else hipLaunchKernelGGL(k_real_win<1>, dim3(wgrid), dim3(256), 0, st, a, rec, soff, wp);
                stage_end(ctx, ST_PAIR);
                if (kind == 0) hipLaunchKernelGGL(k_real_win_final<0>, dim3(fgrid), dim3(256), 0, st, a, wp);
                else hipLaunchKernelGGL(k_real_win_final<1>, dim3(fgrid), dim3(256), 0, st, a, wp);
            } else {
                if (kind == 0) hipLaunchKernelGGL((k_real<0, 1>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
                else hipLaunchKernelGGL((k_real<1, 1>), grid, dim3(256), 0, st, a, spos, q, soff, sbox);
            }
        }
    }
    stage_end(ctx, ST_REAL);
    SE2P_HIP(hipGetLastError());
    if (counting) {
        unsigned long long h[2];
        SE2P_HIP(hipMemcpyAsync(h, p.cnt, sizeof(h), hipMemcpyDeviceToHost, st));
        SE2P_HIP(hipStreamSynchronize(st));
        pair_counts[0] = (int64_t)h[0];
        pair_counts[1] = (int64_t)h[1];
    }
    return 0;
}

int real_sum_multibox(se2p_ctx *ctx, int kind, const se2p_box_desc *boxes, int nbox)
{
    std::vector<RealBox> rb((size_t)nbox);
    for (int i = 0; i < nbox; ++i) {
        const se2p_box_desc &d = boxes[i];
        RealBox &b = rb[(size_t)i];
        b = RealBox{};
        b.psrc = d.psrc;
        b.ptar = d.ptar;
        b.f = d.f;
        b.nv = d.n;
        b.u = d.u;
        b.ns = (int)d.ns;
        b.nt = (int)d.nt;
        b.nsx = d.nside_x;
        b.nsy = d.nside_y;
        b.Lx = d.Lx;
        b.Ly = d.Ly;
        b.xi = d.xi;
    }
    return real_sum_boxes(ctx, kind, rb.data(), nbox, 1, 0, 0, nullptr);
}

static RealBox one_box(const RealGeom &g, const double *f, const double *n, double *u)
{
    RealBox b{};
    b.psrc = g.psrc;
    b.ptar = g.ptar;
    b.f = f;
    b.nv = n;
    b.u = u;
    b.ns = (int)g.ns;
    b.nt = (int)g.nt;
    b.nsx = g.nsx;
    b.nsy = g.nsy;
    b.Lx = g.Lx;
    b.Ly = g.Ly;
    b.xi = g.xi;
    return b;
}

int real_sum(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
             int nrhs, int64_t dstride, int64_t ustride, double *u, int64_t *pair_counts)
{
    RealBox b = one_box(g, f, n, u);
    return real_sum_boxes(ctx, kind, &b, 1, nrhs, dstride, ustride, pair_counts);
}

int real_scratch_bytes(se2p_ctx *ctx, int kind, const RealGeom &g, size_t *bytes)
{
    RealBox b = one_box(g, nullptr, nullptr, nullptr);
    RealPlan p;
    if (int rc = real_plan(ctx, kind, &b, 1, 1, 2 * g.ns, &p)) return rc;
    Carve sizing;
    real_layout(sizing, p);
    *bytes = g.nt > 0 ? sizing.off : 0;
    return 0;
}

int real_sum_fixed(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
                   double *u, char *scratch_base, size_t scratch_bytes)
{
    RealBox b = one_box(g, f, n, u);
    return real_sum_boxes(ctx, kind, &b, 1, 1, 2 * g.ns, 2 * g.nt, nullptr, scratch_base,
                          scratch_bytes);
}

// Adjoint when the targets are the sources: the forward's plan (kept by the
// forward call when it matches, else rebuilt the same way), du gathered into
// its order, one pass of k_real_adj_slp (Stokeslet) or the windowed
// k_real_adj_win (stresslet).
static int real_adjoint_same(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f,
                             const double *n, const double *du, double *dpsrc, double *dptar,
                             double *df, double *dn)
{
    hipStream_t st = ctx->stream;
    if (kind == 0 && upload_table(ctx)) return SE2P_ERR_HIP;
    const bool kept = ctx->kept_real && ctx->kept_real->match(kind, g, f, n);
    RealBox box{};
    box.psrc = g.psrc;
    box.ptar = g.ptar;
    box.f = f;
    box.nv = n;
    box.ns = (int)g.ns;
    box.nt = (int)g.nt;
    box.nsx = g.nsx;
    box.nsy = g.nsy;
    box.Lx = g.Lx;
    box.Ly = g.Ly;
    box.xi = g.xi;
    RealPlan p;
    if (kept) p = ctx->kept_real->plan;
    else if (int rc = real_plan(ctx, kind, &box, 1, 1, 2 * g.ns, &p)) return rc;
    double2 *dut = nullptr;
    // the stresslet's work items: the forward plan's chunk windows (kept, or
    // planned here), its own items, slots and counters
    WinPlan aw{};
    auto layout = [&](Carve &cv) {
        if (!kept) real_layout(cv, p);
        dut = cv.take<double2>(g.nt);
        if (kind == 1) {
            aw.slots = kAdjSlots;
            aw.item_off = cv.take<int>(p.max_chunks);
            aw.slot = cv.take<int>(p.max_chunks);
            aw.item_chunk = cv.take<int>(p.max_chunks + kAdjSlots);
            aw.split = cv.take<int4>(p.max_chunks);
            aw.meta = cv.take<int>(kQueueStride * 9);
            aw.part = cv.take<double>((size_t)kAdjSlots * kAdjNV * 64);
        }
    };
    Carve sizing;
    layout(sizing);
    if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
    Carve cv;
    cv.base = ctx->ws;
    layout(cv);

    stage_begin(ctx, ST_SORT);
    if (!kept)
        if (int rc = real_prepare(ctx, &box, p)) return rc;
    hipLaunchKernelGGL(k_gather_du, dim3((unsigned)((g.nt + 255) / 256)), dim3(256), 0, st, du,
                       p.kss.perm, (int)g.nt, dut);
    stage_end(ctx, ST_SORT);

    const RealBox &B = p.btab.one;
    RealAdjSameArgs a{};
    a.rec = p.rec;
    a.du = dut;
    a.off = p.kss.off;
    a.perm = p.kss.perm;
    a.chunk_off = p.coff;
    a.block_cell = p.bmap;
    a.box = p.sbox;
    a.ncell = (int)p.ncell;
    a.nsx = B.nsx;
    a.nsy = B.nsy;
    a.Lx = B.Lx;
    a.Ly = B.Ly;
    a.xi2 = B.xi2;
    a.cut2 = B.cut2;
    a.self = B.self;
    a.f = f;
    a.nv = n;
    a.dpsrc = dpsrc;
    a.dptar = dptar;
    a.df = df;
    a.dn = dn;
    stage_begin(ctx, ST_REAL);
    if (kind == 0) {
        hipLaunchKernelGGL(k_real_adj_slp, dim3(p.grid8), dim3(256), 0, st, a);
    } else {
        aw.keep = p.win.keep;
        aw.npos = p.win.npos;
        aw.where = p.win.where;
        const unsigned cgrid = (unsigned)((p.max_chunks + kWaves - 1) / kWaves);
        if (!kept) {   // the chunk windows, as the forward plans them
            RealArgs ra{};
            ra.tpos = p.tpos;
            ra.toff = p.kss.off;
            ra.chunk_off = p.coff;
            ra.block_cell = p.bmap;
            ra.ncell = (int)p.ncell;
            ra.boxes = p.btab;
            ra.cell_box = p.cell_box;
            hipLaunchKernelGGL(k_real_win_plan, dim3(cgrid), dim3(256), 0, st, ra, p.kss.off, p.sbox,
                               p.win);
        }
        hipLaunchKernelGGL(k_real_win_scan, dim3(1), dim3(64 * kWinScanWaves), 0, st, p.coff,
                           (int)p.ncell, aw);
        const int64_t items = p.max_chunks + kAdjSlots;
        const int64_t cap = (int64_t)ctx->ncu * (shares_gpu(ctx) ? kAdjBlocksPerCUGroup : kAdjBlocksPerCU);
        const unsigned wgrid = (unsigned)((std::min(items, cap) + 7) & ~int64_t(7));
        stage_begin(ctx, ST_PAIR);
        hipLaunchKernelGGL(k_real_adj_win, dim3(wgrid), dim3(256), 0, st, a, aw);
        stage_end(ctx, ST_PAIR);
        hipLaunchKernelGGL(k_real_adj_win_final, dim3(std::min<unsigned>(cgrid, kWinFinalBlocks)),
                           dim3(256), 0, st, a, aw);
    }
    stage_end(ctx, ST_REAL);
    SE2P_HIP(hipGetLastError());
    return 0;
}

int real_adjoint(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
                 const double *du, double *dpsrc, double *dptar, double *df, double *dn)
{
    hipStream_t st = ctx->stream;
    const int ncell = g.nsx * g.nsy;
    const int sub = 1 << (2 * kSubLog);
    const int nkeys = ncell * sub;
    const int64_t ns = g.ns, nt = g.nt;
    const bool want_src = dpsrc || df || dn;
    // no partners: every gradient is zero
    if (ns == 0 || nt == 0) {
        launch_zero(st, dptar, 2 * nt);
        launch_zero(st, dpsrc, 2 * ns);
        launch_zero(st, df, 2 * ns);
        launch_zero(st, dn, 2 * ns);
        SE2P_HIP(hipGetLastError());
        return 0;
    }
    if (g.ptar == g.psrc && nt == ns)   // targets are the sources
        return real_adjoint_same(ctx, kind, g, f, n, du, dpsrc, dptar, df, dn);
    // separate targets: a target pass and a source pass over their own cell lists
    if (kind == 0 && want_src && upload_table(ctx)) return SE2P_ERR_HIP;
    const int ncomp = kind == 0 ? 2 : 3;
    const int64_t max_t = (nt + 63) / 64 + ncell, max_s = (ns + 63) / 64 + ncell;
    const unsigned grid_t = (unsigned)((max_t + 7) & ~int64_t(7));   // XCD order: multiple of 8
    const unsigned grid_s = (unsigned)((max_s + 7) & ~int64_t(7));

    KeySort kss, kst;
    double2 *spos, *tpos;
    double *q, *dut;
    int *coff_t, *coff_s, *bmap_t, *bmap_s;
    double4 *sbox, *tbox;
    auto layout = [&](Carve &cv) {
        char *p = cv.take<char>(keysort_bytes(ns, nkeys));
        if (p) keysort_carve(p, ns, nkeys, &kss);
        p = cv.take<char>(keysort_bytes(nt, nkeys));
        if (p) keysort_carve(p, nt, nkeys, &kst);
        spos = cv.take<double2>(ns);
        tpos = cv.take<double2>(nt);
        q = cv.take<double>((size_t)ncomp * ns);
        dut = cv.take<double>(2 * (size_t)nt);
        coff_t = cv.take<int>(ncell + 1);
        coff_s = cv.take<int>(ncell + 1);
        bmap_t = cv.take<int>(max_t);
        bmap_s = cv.take<int>(max_s);
        sbox = cv.take<double4>(nkeys);
        tbox = cv.take<double4>(nkeys);
    };
    Carve sizing;
    layout(sizing);
    if (int rc = ensure_workspace(ctx, sizing.off)) return rc;
    Carve cv;
    cv.base = ctx->ws;
    layout(cv);

    stage_begin(ctx, ST_SORT);
    if (int rc = cell_sort(ctx, g.psrc, ns, g.Lx, g.Ly, g.nsx, g.nsy, kSubLog, &kss)) return rc;
    if (int rc = cell_sort(ctx, g.ptar, nt, g.Lx, g.Ly, g.nsx, g.nsy, kSubLog, &kst)) return rc;
    const KeySort &tks = kst;
    const int threads = 256;
    const unsigned sblocks = (unsigned)((ns + threads - 1) / threads);
    const unsigned tblocks = (unsigned)((nt + threads - 1) / threads);
    if (kind == 0)
        hipLaunchKernelGGL(k_gather_sources<0>, dim3(sblocks), dim3(threads), 0, st, g.psrc, f, n,
                           kss.perm, (int)ns, 1, (int64_t)0, spos, q);
    else
        hipLaunchKernelGGL(k_gather_sources<1>, dim3(sblocks), dim3(threads), 0, st, g.psrc, f, n,
                           kss.perm, (int)ns, 1, (int64_t)0, spos, q);
    hipLaunchKernelGGL(k_gather_targets, dim3(tblocks), dim3(threads), 0, st, g.ptar, tks.perm,
                       (int)nt, tpos);
    hipLaunchKernelGGL(k_gather_pairs, dim3(tblocks), dim3(threads), 0, st, du, tks.perm, (int)nt, dut);
    const unsigned kblocks = (unsigned)((nkeys + 3) / 4);   // k_run_bbox: one wavefront per run
    // run boxes of the uniform side: sources for the target pass, targets for
    // the source pass
    if (dptar)
        hipLaunchKernelGGL(k_run_bbox, dim3(kblocks), dim3(threads), 0, st, spos, kss.off, nkeys, sbox);
    if (want_src)
        hipLaunchKernelGGL(k_run_bbox, dim3(kblocks), dim3(threads), 0, st, tpos, tks.off, nkeys, tbox);
    if (dptar) plan_chunks(st, nullptr, tks.off, sub, 64, ncell, coff_t, bmap_t);
    if (want_src) plan_chunks(st, nullptr, kss.off, sub, 64, ncell, coff_s, bmap_s);
    stage_end(ctx, ST_SORT);

    RealAdjArgs a{};
    a.nsx = g.nsx;
    a.nsy = g.nsy;
    a.ncell = ncell;
    a.Lx = g.Lx;
    a.Ly = g.Ly;
    a.xi = g.xi;
    a.xi2 = g.xi * g.xi;
    a.cut2 = g.Lx * g.Ly / g.nsx / g.nsy;   // mex_stokes_slp_real.cpp:127
    a.self = -(0.5 * SE2P_EULER_GAMMA + 1.0) - log(g.xi);
    a.f = f;
    a.nv = n;
    stage_begin(ctx, ST_REAL);
    if (dptar) {
        RealAdjArgs t = a;
        t.lpos = tpos;
        t.lperm = tks.perm;
        t.loff = tks.off;
        t.chunk_off = coff_t;
        t.block_cell = bmap_t;
        t.ldata = dut;
        t.nl = nt;
        t.nu = ns;
        t.dpos = dptar;
        if (kind == 0)
            hipLaunchKernelGGL((k_real_adj<0, 0>), dim3(grid_t), dim3(256), 0, st, t, spos, q, kss.off, sbox);
        else
            hipLaunchKernelGGL((k_real_adj<1, 0>), dim3(grid_t), dim3(256), 0, st, t, spos, q, kss.off, sbox);
    }
    if (dpsrc || df || dn) {
        RealAdjArgs s = a;
        s.lpos = spos;
        s.lperm = kss.perm;
        s.loff = kss.off;
        s.chunk_off = coff_s;
        s.block_cell = bmap_s;
        s.ldata = q;
        s.nl = ns;
        s.nu = nt;
        s.dpos = dpsrc;
        s.df = df;
        s.dn = dn;
        if (kind == 0)
            hipLaunchKernelGGL((k_real_adj<0, 1>), dim3(grid_s), dim3(256), 0, st, s, tpos, dut, tks.off, tbox);
        else
            hipLaunchKernelGGL((k_real_adj<1, 1>), dim3(grid_s), dim3(256), 0, st, s, tpos, dut, tks.off, tbox);
    }
    stage_end(ctx, ST_REAL);
    SE2P_HIP(hipGetLastError());
    return 0;
}

}  // namespace se2p
