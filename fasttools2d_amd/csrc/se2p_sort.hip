// se2p_sort.hip -- stable key sort of point sets (cell lists and gridding
// bins).
//
// The reference builds its cell lists with a serial counting sort that keeps
// points of a box in input order (ewald_tools.cpp:7-80).  Here the same order
// comes from a stable LSD radix sort of (key, index) pairs (hipCUB), followed
// by a lower-bound search per key for the run offsets.  Deterministic: no
// atomics decide any position.
#include <hipcub/hipcub.hpp>

#include "se2p_device.h"
#include "se2p_internal.h"

namespace se2p {

static int key_bits(int nkeys)
{
    int b = 1;
    while ((1LL << b) < (long long)nkeys) ++b;
    return b;
}

size_t keysort_bytes(int64_t n, int nkeys, size_t *temp_bytes)
{
    size_t tb = 0;
    int nn = (int)(n > 0 ? n : 1);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const int *)nullptr, (int *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, nn, 0,
                                             key_bits(nkeys));
    *temp_bytes = tb;
    size_t a = 0;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    a += 4 * up(sizeof(int) * (size_t)nn);          // keys_in, keys_out, idx_in, perm
    a += up(sizeof(int) * ((size_t)nkeys + 1));     // off
    a += up(tb);
    return a;
}

void keysort_carve(char *base, int64_t n, int nkeys, size_t temp_bytes, KeySort *ks)
{
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t nn = (size_t)(n > 0 ? n : 1);
    size_t o = 0;
    ks->keys_in = (int *)(base + o);  o += up(sizeof(int) * nn);
    ks->keys_out = (int *)(base + o); o += up(sizeof(int) * nn);
    ks->idx_in = (int *)(base + o);   o += up(sizeof(int) * nn);
    ks->perm = (int *)(base + o);     o += up(sizeof(int) * nn);
    ks->off = (int *)(base + o);      o += up(sizeof(int) * ((size_t)nkeys + 1));
    ks->temp = base + o;
    ks->temp_bytes = temp_bytes;
}

// off[k] = first position whose sorted key is >= k (k = 0..nkeys).
__global__ void k_run_offsets(const int *__restrict__ keys, int n, int nkeys, int *__restrict__ off)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    off[k] = lo;
}

int keysort_run(se2p_ctx *ctx, KeySort *ks, int64_t n, int nkeys)
{
    hipStream_t st = ctx->stream;
    if (n > 0) {
        size_t tb = ks->temp_bytes;
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(ks->temp, tb, ks->keys_in, ks->keys_out,
                                                          ks->idx_in, ks->perm, (int)n, 0,
                                                          key_bits(nkeys), st);
        if (e != hipSuccess)
            return fail(ctx, SE2P_ERR_HIP, "radix sort failed: %s", hipGetErrorString(e));
    }
    int threads = 256;
    int blocks = (nkeys + 1 + threads - 1) / threads;
    hipLaunchKernelGGL(k_run_offsets, dim3(blocks), dim3(threads), 0, st, ks->keys_out, (int)n,
                       nkeys, ks->off);
    return 0;
}

__global__ void k_cell_keys(const double *__restrict__ p, int n, double Lx, double Ly, int nsx,
                            int nsy, int *__restrict__ keys, int *__restrict__ idx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = se2p_cell_of(p[2 * i], p[2 * i + 1], Lx, Ly, nsx, nsy);
    idx[i] = i;
}

void launch_cell_keys(hipStream_t st, const double *p, int64_t n, double Lx, double Ly, int nsx,
                      int nsy, int *keys, int *idx)
{
    if (n <= 0) return;
    int threads = 256;
    int blocks = (int)((n + threads - 1) / threads);
    hipLaunchKernelGGL(k_cell_keys, dim3(blocks), dim3(threads), 0, st, p, (int)n, Lx, Ly, nsx,
                       nsy, keys, idx);
}

}  // namespace se2p
