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
// GPU mapping: targets and sources are sorted by cell (stable, so a cell's
// points keep input order).  Each 64-lane wavefront takes one chunk of up to
// 64 targets of one cell and walks the 9 source cells in the reference's
// order; every lane keeps its own target's sum in registers.  Source data is
// wave-uniform, so it arrives through the scalar cache.  E1 comes from a
// piecewise polynomial table staged in LDS.  No atomics: each output element
// is written once, so results are deterministic.
#include <hipcub/hipcub.hpp>

#include "se2p_device.h"
#include "se2p_internal.h"

namespace se2p {

static __constant__ double c_e1_table[SE2P_E1_NPIECE * SE2P_E1_ROW];

// Neighbour j = 0..7 in the reference's order (mex_stokes_slp_real.cpp:61-62):
// the 3x3 block in x-major order with the centre left out.
__device__ __forceinline__ int nb_dx(int j) { return (j < 4 ? j : j + 1) / 3 - 1; }
__device__ __forceinline__ int nb_dy(int j) { return (j < 4 ? j : j + 1) % 3 - 1; }

struct RealArgs {
    const double2 *tpos;   // targets in cell order
    const int *tperm;      // cell order -> caller's index
    const int *toff;       // [ncell+1]
    const int *chunk_off;  // [ncell+1] exclusive scan of ceil(targets/64)
    const double2 *spos;   // sources in cell order
    const double *q;       // densities in cell order, [nrhs][ncomp][ns]
    const int *soff;       // [ncell+1]
    int64_t ns;
    int nsx, nsy, ncell, nrhs;
    double Lx, Ly, xi, xi2, cut2, self;
    double *u;             // [nrhs] x (2 x nt), interleaved, caller's order
    int64_t ustride;
    unsigned long long *pair_counts;   // [2] when counting, else null
};

// ---- pair kernels --------------------------------------------------------
// Stokeslet pair, before 1/4pi: 0.5 E1(xi^2 r^2) f + e^{-xi^2 r^2}((d.f) d/r^2 - f).
// Written as in mex_stokes_slp_real.cpp:207-225: c = d1 f2 - d2 f1,
// u1 += E/2 f1 + (e/r2) d2 c,  u2 += E/2 f2 - (e/r2) d1 c.
template <int RB>
__device__ __forceinline__ void slp_pair(double d1, double d2, double r2, const RealArgs &a,
                                         const double *e1tab, const double *q, int64_t qs,
                                         double (&acc)[RB][2])
{
    double x = r2 * a.xi * a.xi;
    double e = exp(-x);
    double E = se2p_e1(x, e, e1tab);
    double g = e / r2;
    double hE = 0.5 * E;
    double g1 = g * d1, g2 = g * d2;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        double f1 = q[(2 * r) * qs], f2 = q[(2 * r + 1) * qs];
        double c = d1 * f2 - d2 * f1;
        acc[r][0] += hE * f1 + g2 * c;
        acc[r][1] += hE * f2 - g1 * c;
    }
}

// Stresslet pair, before 1/4pi (mex_stokes_dlp_real.cpp:125-138).  q holds
// (S11, S12+S21, S22) with S = f (x) n.
template <int RB>
__device__ __forceinline__ void dlp_pair(double d1, double d2, double r2, const RealArgs &a,
                                         const double *q, int64_t qs, double (&acc)[RB][2])
{
    double x = a.xi2 * r2;
    double e2 = exp(-x);
    double ir2 = 1.0 / r2;
    double fb = -4.0 * (1.0 + x) * ir2 * ir2;
    double pre = 2.0 * a.xi2;
    double d11 = d1 * d1, d12 = d1 * d2, d22 = d2 * d2;
    double t111 = e2 * (d11 * d1 * fb + pre * 3.0 * d1);
    double t112 = e2 * (d11 * d2 * fb + pre * d2);
    double t122 = e2 * (d12 * d2 * fb + pre * d1);
    double t222 = e2 * (d22 * d2 * fb + pre * 3.0 * d2);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        double s0 = q[(3 * r) * qs], s1 = q[(3 * r + 1) * qs], s2 = q[(3 * r + 2) * qs];
        acc[r][0] += t111 * s0 + t112 * s1 + t122 * s2;
        acc[r][1] += t112 * s0 + t122 * s1 + t222 * s2;
    }
}

// ---- the cell-list kernel ------------------------------------------------
template <int KIND, int RB, bool COUNT>
__global__ __launch_bounds__(64) void k_real(RealArgs a)
{
    __shared__ double e1tab[KIND == 0 ? SE2P_E1_NPIECE * SE2P_E1_ROW : 1];
    const int lane = threadIdx.x;
    if (KIND == 0) {
        for (int i = lane; i < SE2P_E1_NPIECE * SE2P_E1_ROW; i += 64) e1tab[i] = c_e1_table[i];
        __syncthreads();
    }

    // block -> (cell, chunk) by upper bound in chunk_off
    const int b = blockIdx.x;
    int lo = 0, hi = a.ncell;   // find largest c with chunk_off[c] <= b
    if (b >= a.chunk_off[a.ncell]) return;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (a.chunk_off[mid] <= b) lo = mid;
        else hi = mid;
    }
    const int cell = lo;
    const int t0 = a.toff[cell] + 64 * (b - a.chunk_off[cell]);
    const int t1 = a.toff[cell + 1];
    const int t = t0 + lane;
    const bool active = t < t1;
    double tx = 0.0, ty = 0.0;
    if (active) {
        double2 p = a.tpos[t];
        tx = p.x;
        ty = p.y;
    } else {
        tx = ty = __builtin_nan("");   // inactive lanes compute garbage, never stored
    }

    const int ncomp = KIND == 0 ? 2 : 3;
    const int64_t qs = a.ns;
    unsigned long long n_in = 0, n_out = 0;

    for (int r0 = 0; r0 < a.nrhs; r0 += RB) {
        double acc[RB][2];
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r][0] = acc[r][1] = 0.0;
        const double *qb = a.q + (int64_t)r0 * ncomp * qs;

        // own cell: every pair, no cutoff
        {
            const int s0 = a.soff[cell], s1 = a.soff[cell + 1];
            for (int s = s0; s < s1; ++s) {
                const double2 sp = a.spos[s];
                double d1, d2;
                if (KIND == 0) {   // reference: target - source
                    d1 = tx - sp.x;
                    d2 = ty - sp.y;
                } else {           // reference: source - target
                    d1 = sp.x - tx;
                    d2 = sp.y - ty;
                }
                double r2 = se2p_sqsum(d1, d2);
                if (COUNT) {
                    n_in += (active && r2 != 0.0) ? 1 : 0;
                    continue;
                }
                if (r2 == 0.0) {
                    if (KIND == 0) {
#pragma unroll
                        for (int r = 0; r < RB; ++r) {
                            acc[r][0] += a.self * qb[(2 * r) * qs + s];
                            acc[r][1] += a.self * qb[(2 * r + 1) * qs + s];
                        }
                    }
                } else {
                    if (KIND == 0) slp_pair<RB>(d1, d2, r2, a, e1tab, qb + s, qs, acc);
                    else dlp_pair<RB>(d1, d2, r2, a, qb + s, qs, acc);
                }
            }
        }
        // eight neighbour cells, periodic offsets, r^2 < cut2
        const int cx = cell % a.nsx, cy = cell / a.nsx;
#pragma unroll 1
        for (int j = 0; j < 8; ++j) {
            const int qx = cx + nb_dx(j), qy = cy + nb_dy(j);
            const int wx = (qx + a.nsx) % a.nsx, wy = (qy + a.nsy) % a.nsy;
            const int sb = wy * a.nsx + wx;
            const int s0 = a.soff[sb], s1 = a.soff[sb + 1];
            if (s0 == s1) continue;
            const double zx = (a.Lx * (qx - wx)) / a.nsx;
            const double zy = (a.Ly * (qy - wy)) / a.nsy;
            // reference SLP: d = src - (tgt - z); DLP: d = (src - tgt) + z
            const double sx = tx - zx, sy = ty - zy;
            for (int s = s0; s < s1; ++s) {
                const double2 sp = a.spos[s];
                double d1, d2;
                if (KIND == 0) {
                    d1 = sp.x - sx;
                    d2 = sp.y - sy;
                } else {
                    d1 = (sp.x - tx) + zx;
                    d2 = (sp.y - ty) + zy;
                }
                double r2 = se2p_sqsum(d1, d2);
                if (COUNT) {
                    if (active) {
                        if (r2 < a.cut2) ++n_in;
                        else ++n_out;
                    }
                    continue;
                }
                if (r2 < a.cut2) {
                    if (KIND == 0) slp_pair<RB>(d1, d2, r2, a, e1tab, qb + s, qs, acc);
                    else dlp_pair<RB>(d1, d2, r2, a, qb + s, qs, acc);
                }
            }
        }
        if (!COUNT && active) {
            const int o = a.tperm[t];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                double *ur = a.u + (int64_t)(r0 + r) * a.ustride;
                if (KIND == 0) {
                    ur[2 * o] = acc[r][0] / (4 * SE2P_PI);
                    ur[2 * o + 1] = acc[r][1] / (4 * SE2P_PI);
                } else {
                    ur[2 * o] = acc[r][0] / 4 / SE2P_PI;
                    ur[2 * o + 1] = acc[r][1] / 4 / SE2P_PI;
                }
            }
        }
        if (COUNT) break;
    }
    if (COUNT) {
        // wave reduction then one atomic per wave (integer: deterministic)
        for (int off = 32; off > 0; off >>= 1) {
            n_in += __shfl_down(n_in, off);
            n_out += __shfl_down(n_out, off);
        }
        if (lane == 0) {
            atomicAdd(&a.pair_counts[0], n_in);
            atomicAdd(&a.pair_counts[1], n_out);
        }
    }
}

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

// chunk counts per cell, then an exclusive scan
__global__ void k_chunk_counts(const int *__restrict__ toff, int ncell, int *__restrict__ cnt)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    cnt[c] = c < ncell ? (toff[c + 1] - toff[c] + 63) / 64 : 0;
}

static bool table_uploaded[64] = {};

static int upload_table(se2p_ctx *ctx)
{
    if (ctx->device >= 0 && ctx->device < 64 && table_uploaded[ctx->device]) return 0;
    static double host[SE2P_E1_NPIECE * SE2P_E1_ROW];
    for (int p = 0; p < SE2P_E1_NPIECE; ++p) {
        double *row = host + p * SE2P_E1_ROW;
        row[0] = se2p_e1_scale_shift[p][0];
        row[1] = se2p_e1_scale_shift[p][1];
        for (int k = 0; k < SE2P_E1_NCOEF; ++k) row[2 + k] = se2p_e1_coef[p][k];
        for (int k = 2 + SE2P_E1_NCOEF; k < SE2P_E1_ROW; ++k) row[k] = 0.0;
    }
    SE2P_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_e1_table), host, sizeof(host)));
    if (ctx->device >= 0 && ctx->device < 64) table_uploaded[ctx->device] = true;
    return 0;
}

// ---- host driver ----------------------------------------------------------
int real_sum(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
             int nrhs, int64_t dstride, int64_t ustride, double *u, int64_t *pair_counts)
{
    hipStream_t st = ctx->stream;
    const int ncell = g.nsx * g.nsy;
    const int64_t ns = g.ns, nt = g.nt;
    const bool counting = pair_counts != nullptr;
    if (nt == 0) return 0;
    if (kind == 0 && upload_table(ctx)) return SE2P_ERR_HIP;

    // workspace layout
    size_t tb_s = 0, tb_t = 0;
    size_t ks_s = keysort_bytes(ns, ncell, &tb_s);
    size_t ks_t = keysort_bytes(nt, ncell, &tb_t);
    const int ncomp = kind == 0 ? 2 : 3;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t scan_tb = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tb, (int *)nullptr, (int *)nullptr,
                                           ncell + 1);
    size_t need = ks_s + ks_t + up(sizeof(double2) * (size_t)(ns > 0 ? ns : 1)) +
                  up(sizeof(double2) * (size_t)nt) +
                  up(sizeof(double) * (size_t)ncomp * nrhs * (size_t)(ns > 0 ? ns : 1)) +
                  2 * up(sizeof(int) * (size_t)(ncell + 1)) + up(scan_tb) +
                  up(2 * sizeof(unsigned long long));
    if (int rc = ensure_workspace(ctx, need)) return rc;
    char *base = ctx->ws;
    size_t o = 0;
    KeySort kss, kst;
    keysort_carve(base + o, ns, ncell, tb_s, &kss); o += ks_s;
    keysort_carve(base + o, nt, ncell, tb_t, &kst); o += ks_t;
    double2 *spos = (double2 *)(base + o); o += up(sizeof(double2) * (size_t)(ns > 0 ? ns : 1));
    double2 *tpos = (double2 *)(base + o); o += up(sizeof(double2) * (size_t)nt);
    double *q = (double *)(base + o); o += up(sizeof(double) * (size_t)ncomp * nrhs * (size_t)(ns > 0 ? ns : 1));
    int *ccnt = (int *)(base + o); o += up(sizeof(int) * (size_t)(ncell + 1));
    int *coff = (int *)(base + o); o += up(sizeof(int) * (size_t)(ncell + 1));
    void *scan_tmp = base + o; o += up(scan_tb);
    unsigned long long *cnt = (unsigned long long *)(base + o); o += up(2 * sizeof(unsigned long long));

    stage_begin(ctx, ST_SORT);
    launch_cell_keys(st, g.psrc, ns, g.Lx, g.Ly, g.nsx, g.nsy, kss.keys_in, kss.idx_in);
    if (int rc = keysort_run(ctx, &kss, ns, ncell)) return rc;
    launch_cell_keys(st, g.ptar, nt, g.Lx, g.Ly, g.nsx, g.nsy, kst.keys_in, kst.idx_in);
    if (int rc = keysort_run(ctx, &kst, nt, ncell)) return rc;
    const int threads = 256;
    if (ns > 0) {
        int blocks = (int)((ns + threads - 1) / threads);
        if (kind == 0)
            hipLaunchKernelGGL(k_gather_sources<0>, dim3(blocks), dim3(threads), 0, st, g.psrc, f,
                               n, kss.perm, (int)ns, nrhs, dstride, spos, q);
        else
            hipLaunchKernelGGL(k_gather_sources<1>, dim3(blocks), dim3(threads), 0, st, g.psrc, f,
                               n, kss.perm, (int)ns, nrhs, dstride, spos, q);
    }
    {
        int blocks = (int)((nt + threads - 1) / threads);
        hipLaunchKernelGGL(k_gather_targets, dim3(blocks), dim3(threads), 0, st, g.ptar, kst.perm,
                           (int)nt, tpos);
        hipLaunchKernelGGL(k_chunk_counts, dim3((ncell + 1 + threads - 1) / threads),
                           dim3(threads), 0, st, kst.off, ncell, ccnt);
        size_t tb = scan_tb;
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, ccnt, coff, ncell + 1, st);
        if (e != hipSuccess) return fail(ctx, SE2P_ERR_HIP, "scan failed: %s", hipGetErrorString(e));
    }
    stage_end(ctx, ST_SORT);

    RealArgs a;
    a.tpos = tpos;
    a.tperm = kst.perm;
    a.toff = kst.off;
    a.chunk_off = coff;
    a.spos = spos;
    a.q = q;
    a.soff = kss.off;
    a.ns = ns;
    a.nsx = g.nsx;
    a.nsy = g.nsy;
    a.ncell = ncell;
    a.nrhs = nrhs;
    a.Lx = g.Lx;
    a.Ly = g.Ly;
    a.xi = g.xi;
    a.xi2 = g.xi * g.xi;
    a.cut2 = g.Lx * g.Ly / g.nsx / g.nsy;   // mex_stokes_slp_real.cpp:127
    a.self = -(0.5 * SE2P_EULER_GAMMA + 1.0) - log(g.xi);
    a.u = u;
    a.ustride = ustride;
    a.pair_counts = cnt;

    // upper bound on chunks: every non-empty cell wastes < 1 chunk
    const int64_t max_chunks = (nt + 63) / 64 + ncell;
    stage_begin(ctx, ST_REAL);
    if (counting) {
        SE2P_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), st));
        if (kind == 0) hipLaunchKernelGGL((k_real<0, 1, true>), dim3(max_chunks), dim3(64), 0, st, a);
        else hipLaunchKernelGGL((k_real<1, 1, true>), dim3(max_chunks), dim3(64), 0, st, a);
    } else if (kind == 0) {
        if (nrhs % 4 == 0)
            hipLaunchKernelGGL((k_real<0, 4, false>), dim3(max_chunks), dim3(64), 0, st, a);
        else
            hipLaunchKernelGGL((k_real<0, 1, false>), dim3(max_chunks), dim3(64), 0, st, a);
    } else {
        if (nrhs % 4 == 0)
            hipLaunchKernelGGL((k_real<1, 4, false>), dim3(max_chunks), dim3(64), 0, st, a);
        else
            hipLaunchKernelGGL((k_real<1, 1, false>), dim3(max_chunks), dim3(64), 0, st, a);
    }
    stage_end(ctx, ST_REAL);
    SE2P_HIP(hipGetLastError());
    if (counting) {
        unsigned long long h[2];
        SE2P_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
        SE2P_HIP(hipStreamSynchronize(st));
        pair_counts[0] = (int64_t)h[0];
        pair_counts[1] = (int64_t)h[1];
    }
    return 0;
}

}  // namespace se2p
