// se2p_compose.hip -- small kernels of the driver entry points: wrapping into
// the primary box, deterministic reductions (sum of f^2 for the error
// estimates, the stresslet's k = 0 mode) and the final composition.
#include "se2p_device.h"
#include "se2p_internal.h"

namespace se2p {

// mod(x + L/2, L) - L/2 with mod(a, L) = a - floor(a/L) L
// (StokesSLP_ewald_2p.m:75-78, StokesDLP_ewald_2p.m:94-97)
__global__ void k_wrap(const double *__restrict__ p, int n, double Lx, double Ly,
                       double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double tx = p[2 * i] + Lx / 2, ty = p[2 * i + 1] + Ly / 2;
    out[2 * i] = (tx - floor(tx / Lx) * Lx) - Lx / 2;
    out[2 * i + 1] = (ty - floor(ty / Ly) * Ly) - Ly / 2;
}

// Fixed-order block reduction of 256 values (tree in LDS).
__device__ inline double block_sum_256(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partial[b*3 + {0,1,2}] over block b's points: sum f.f, sum (f.n) x, sum (f.n) y
__global__ __launch_bounds__(256) void k_driver_partials(const double *__restrict__ f,
                                                         const double *__restrict__ nv,
                                                         const double *__restrict__ x, int n,
                                                         double *__restrict__ partial)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ff = 0.0, zx = 0.0, zy = 0.0;
    if (i < n) {
        const double f1 = f[2 * i], f2 = f[2 * i + 1];
        ff = f1 * f1 + f2 * f2;
        if (nv) {
            const double fn = f1 * nv[2 * i] + f2 * nv[2 * i + 1];
            zx = fn * x[2 * i];
            zy = fn * x[2 * i + 1];
        }
    }
    ff = block_sum_256(ff, sh);
    zx = block_sum_256(zx, sh);
    zy = block_sum_256(zy, sh);
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = ff;
        partial[3 * blockIdx.x + 1] = zx;
        partial[3 * blockIdx.x + 2] = zy;
    }
}

// out[0..2] = sums of the partials, in block order (one workgroup)
__global__ __launch_bounds__(256) void k_driver_total(const double *__restrict__ partial, int nb,
                                                      double *__restrict__ out)
{
    __shared__ double sh[256];
    for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) acc += partial[3 * b + c];
        acc = block_sum_256(acc, sh);
        if (threadIdx.x == 0) out[c] = acc;
    }
}

// SLP: u = ur + uk (StokesSLP_ewald_2p.m:133).  DLP: uk -= zero mode, then
// u = -(ur + uk) (StokesDLP_ewald_2p.m:150-151,158); uk is updated in place.
__global__ void k_compose(const double *__restrict__ ur, double *__restrict__ uk, int nt, int dlp,
                          const double *__restrict__ zm, double LxLy, double *__restrict__ u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double k1 = uk[2 * i], k2 = uk[2 * i + 1];
    if (dlp) {
        k1 = k1 - zm[1] / LxLy;
        k2 = k2 - zm[2] / LxLy;
        uk[2 * i] = k1;
        uk[2 * i + 1] = k2;
        u[2 * i] = -(ur[2 * i] + k1);
        u[2 * i + 1] = -(ur[2 * i + 1] + k2);
    } else {
        u[2 * i] = ur[2 * i] + k1;
        u[2 * i + 1] = ur[2 * i + 1] + k2;
    }
}

__global__ void k_fill_zero(double *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

void launch_zero(hipStream_t st, double *p, int64_t n)
{
    if (p == nullptr || n <= 0) return;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

void launch_wrap(hipStream_t st, const double *p, int64_t n, double Lx, double Ly, double *out)
{
    if (n > 0)
        hipLaunchKernelGGL(k_wrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, (int)n, Lx,
                           Ly, out);
}

void launch_driver_sums(hipStream_t st, const double *f, const double *nv, const double *x,
                        int64_t n, double *partial, double *out)
{
    const int nb = (int)((n + 255) / 256);
    if (nb > 0)
        hipLaunchKernelGGL(k_driver_partials, dim3(nb), dim3(256), 0, st, f, nv, x, (int)n, partial);
    hipLaunchKernelGGL(k_driver_total, dim3(1), dim3(256), 0, st, partial, nb, out);
}

void launch_compose(hipStream_t st, const double *ur, double *uk, int64_t nt, int dlp,
                    const double *zm, double LxLy, double *u)
{
    if (nt > 0)
        hipLaunchKernelGGL(k_compose, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, ur, uk,
                           (int)nt, dlp, zm, LxLy, u);
}

}  // namespace se2p
