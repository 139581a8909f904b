// se2p_sort.hip -- counting sort of point sets by an integer key (cell lists
// and gridding bins).
//
// The reference builds its cell lists with a serial, stable counting sort
// (ewald_tools.cpp:7-80): points of a box keep their input order.  Here:
//   1. key per point, slot = atomicAdd(count[key])      (slot order arbitrary)
//   2. exclusive scan of the counts                      (one workgroup)
//   3. scatter i to off[key] + slot
//   4. (stable sorts only) rank each element inside its key run by input
//      index and scatter again, which restores input order within a run.
// Step 4 makes the order -- and hence every floating-point summation order
// downstream -- independent of atomic timing.  Sorts whose order cannot
// change any result (targets of the real-space sum) skip it.
#include "se2p_device.h"
#include "se2p_internal.h"

namespace se2p {

static size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

size_t keysort_bytes(int64_t n, int nkeys)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    return 4 * up256(sizeof(int) * nn) + 2 * up256(sizeof(int) * ((size_t)nkeys + 1));
}

void keysort_carve(char *base, int64_t n, int nkeys, KeySort *ks)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    size_t o = 0;
    ks->key = (int *)(base + o);   o += up256(sizeof(int) * nn);
    ks->slot = (int *)(base + o);  o += up256(sizeof(int) * nn);
    ks->tmp = (int *)(base + o);   o += up256(sizeof(int) * nn);
    ks->perm = (int *)(base + o);  o += up256(sizeof(int) * nn);
    ks->count = (int *)(base + o); o += up256(sizeof(int) * ((size_t)nkeys + 1));
    ks->off = (int *)(base + o);   o += up256(sizeof(int) * ((size_t)nkeys + 1));
}

__global__ void k_fill_int(int *__restrict__ a, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Exclusive scan of a[0..n) into out[0..n] (out[n] = total), one workgroup.
__global__ __launch_bounds__(1024) void k_scan_one_block(const int *__restrict__ a, int n,
                                                         int *__restrict__ out)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partials
    for (int d = 1; d < 1024; d <<= 1) {
        int v = (t >= d) ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = (t > 0) ? part[t - 1] : 0;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += a[i];
    }
    if (t == 1023) out[n] = part[1023];
}

__global__ void k_scatter(const int *__restrict__ key, const int *__restrict__ slot,
                          const int *__restrict__ off, int n, int *__restrict__ dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[off[key[i]] + slot[i]] = i;
}

// Stable order inside each key run: rank = number of run members with a
// smaller input index.
__global__ void k_rank_in_run(const int *__restrict__ key, const int *__restrict__ tmp,
                              const int *__restrict__ off, int n, int *__restrict__ perm)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int e = tmp[p];
    const int k = key[e];
    const int s0 = off[k], s1 = off[k + 1];
    int r = 0;
    for (int q = s0; q < s1; ++q) r += (tmp[q] < e) ? 1 : 0;
    perm[s0 + r] = e;
}

void scan_ints(hipStream_t st, const int *a, int n, int *out)
{
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, st, a, n, out);
}

int keysort_finish(se2p_ctx *ctx, KeySort *ks, int64_t n, int nkeys, bool stable)
{
    hipStream_t st = ctx->stream;
    const int threads = 256;
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, st, ks->count, nkeys, ks->off);
    if (n > 0) {
        const unsigned blocks = (unsigned)((n + threads - 1) / threads);
        hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(threads), 0, st, ks->key, ks->slot,
                           ks->off, (int)n, stable ? ks->tmp : ks->perm);
        if (stable)
            hipLaunchKernelGGL(k_rank_in_run, dim3(blocks), dim3(threads), 0, st, ks->key,
                               ks->tmp, ks->off, (int)n, ks->perm);
    }
    return 0;
}

void keysort_clear(hipStream_t st, KeySort *ks, int nkeys)
{
    hipLaunchKernelGGL(k_fill_int, dim3((nkeys + 1 + 255) / 256), dim3(256), 0, st, ks->count,
                       nkeys + 1, 0);
}

// cell key (ewald_tools.cpp:28-35); with sub > 0 the key also carries a
// sub x sub Morton-ordered sub-cell so that consecutive targets of a cell are
// spatially close.
__global__ void k_cell_keys(const double *__restrict__ p, int n, double Lx, double Ly, int nsx,
                            int nsy, int sublog, int *__restrict__ key, int *__restrict__ slot,
                            int *__restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = p[2 * i], y = p[2 * i + 1];
    int k = se2p_cell_of(x, y, Lx, Ly, nsx, nsy);
    if (sublog > 0) {
        const int sub = 1 << sublog;
        const int bx = k % nsx, by = k / nsx;
        int sx = (int)floor((nsx * (x / Lx + 0.5) - bx) * sub);
        int sy = (int)floor((nsy * (y / Ly + 0.5) - by) * sub);
        sx = sx < 0 ? 0 : (sx >= sub ? sub - 1 : sx);
        sy = sy < 0 ? 0 : (sy >= sub ? sub - 1 : sy);
        int m = 0;
        for (int b = 0; b < sublog; ++b)
            m |= (((sx >> b) & 1) << (2 * b)) | (((sy >> b) & 1) << (2 * b + 1));
        k = (k << (2 * sublog)) | m;
    }
    key[i] = k;
    slot[i] = atomicAdd(&count[k], 1);
}

int cell_sort(se2p_ctx *ctx, const double *p, int64_t n, double Lx, double Ly, int nsx, int nsy,
              int sublog, bool stable, KeySort *ks)
{
    hipStream_t st = ctx->stream;
    const int nkeys = nsx * nsy << (2 * sublog);
    keysort_clear(st, ks, nkeys);
    if (n > 0)
        hipLaunchKernelGGL(k_cell_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p,
                           (int)n, Lx, Ly, nsx, nsy, sublog, ks->key, ks->slot, ks->count);
    return keysort_finish(ctx, ks, n, nkeys, stable);
}

}  // namespace se2p
