// se2p_sort.hip -- stable sort of point sets by an integer key (cell lists
// and gridding bins).
//
// The reference builds its cell lists with a serial, stable counting sort
// (ewald_tools.cpp:7-80): points of a box keep their input order.  Here a
// least-significant-digit radix sort with 8-bit digits does the same job on
// the GPU and is stable by construction, so the order -- and every
// floating-point summation order downstream -- is reproducible run to run.
// One pass is three launches:
//   hist    : per-block digit histogram (LDS counters), stored digit-major;
//   scan    : exclusive scan of the histograms -> global digit/block offsets;
//   scatter : each block ranks its elements stably (wavefront ballots decide
//             equal digits, LDS prefix over the four wavefronts and over the
//             block's four rounds) and writes them to their final slots.
// A final launch turns the sorted keys into run offsets (lower bound per key).
#include "se2p_device.h"
#include "se2p_fsort.h"
#include "se2p_internal.h"

namespace se2p {

constexpr int kRadixBlock = 1024;   // elements per block (4 rounds of 256)

static size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

static int key_bits(int nkeys)
{
    int b = 1;
    while ((1LL << b) < (long long)nkeys) ++b;
    return b;
}

static int radix_blocks(int64_t n) { return (int)((n + kRadixBlock - 1) / kRadixBlock); }

size_t keysort_bytes(int64_t n, int nkeys)
{
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t nb = (size_t)(radix_blocks(n) > 0 ? radix_blocks(n) : 1);
    return 4 * up256(sizeof(int) * nn) + up256(sizeof(int) * ((size_t)nkeys + 1)) +
           2 * up256(sizeof(int) * (256 * nb + 256));
}

void keysort_carve(char *base, int64_t n, int nkeys, KeySort *ks)
{
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t nb = (size_t)(radix_blocks(n) > 0 ? radix_blocks(n) : 1);
    size_t o = 0;
    ks->key = (int *)(base + o);    o += up256(sizeof(int) * nn);
    ks->key2 = (int *)(base + o);   o += up256(sizeof(int) * nn);
    ks->tmp = (int *)(base + o);    o += up256(sizeof(int) * nn);
    ks->perm = (int *)(base + o);   o += up256(sizeof(int) * nn);
    ks->off = (int *)(base + o);    o += up256(sizeof(int) * ((size_t)nkeys + 1));
    ks->hist = (int *)(base + o);   o += up256(sizeof(int) * (256 * nb + 256));   // + digit totals
    ks->hoff = (int *)(base + o);   o += up256(sizeof(int) * (256 * nb + 256));
}

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the exclusive prefix and writes the block total to *total.
__device__ inline int block_exclusive_scan_256(int v, int *wsum, int *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// Exclusive scan of a[0..n) into out[0..n] (out[n] = total), one 256-thread
// workgroup walking the array in coalesced chunks of 256 with a carry.
__global__ __launch_bounds__(256) void k_scan_one_block(const int *__restrict__ a, int n,
                                                        int *__restrict__ out)
{
    __shared__ int wsum[4];
    int carry = 0;
    for (int b = 0; b < n; b += 256) {
        const int i = b + threadIdx.x;
        const int v = i < n ? a[i] : 0;
        int tot;
        const int ex = block_exclusive_scan_256(v, wsum, &tot);
        if (i < n) out[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

void scan_ints(hipStream_t st, const int *a, int n, int *out)
{
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(256), 0, st, a, n, out);
}

__global__ __launch_bounds__(256) void k_radix_hist(const int *__restrict__ key, int n, int shift,
                                                    int nblocks, int *__restrict__ hist)
{
    __shared__ int cnt[256];
    const int t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const int b = blockIdx.x;
    for (int j = 0; j < kRadixBlock / 256; ++j) {
        const int i = b * kRadixBlock + j * 256 + t;
        if (i < n) atomicAdd(&cnt[(key[i] >> shift) & 255], 1);   // integer counts
    }
    __syncthreads();
    hist[t * nblocks + b] = cnt[t];
}

// Per digit d (one workgroup each): exclusive scan of hist[d][0..nb) into
// hoff[d][...] and the digit total into dtot[d].
__global__ __launch_bounds__(256) void k_radix_digit_scan(const int *__restrict__ hist, int nb,
                                                          int *__restrict__ hoff,
                                                          int *__restrict__ dtot)
{
    __shared__ int wsum[4];
    const int d = blockIdx.x;
    int carry = 0;
    for (int b = 0; b < nb; b += 256) {
        const int i = b + threadIdx.x;
        const int v = i < nb ? hist[d * nb + i] : 0;
        int tot;
        const int ex = block_exclusive_scan_256(v, wsum, &tot);
        if (i < nb) hoff[d * nb + i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) dtot[d] = carry;
}

// Stable scatter of one radix pass.  vin == nullptr means the identity
// (original indices).
__global__ __launch_bounds__(256) void k_radix_scatter(const int *__restrict__ kin,
                                                       const int *__restrict__ vin, int n,
                                                       int shift, int nblocks,
                                                       const int *__restrict__ hoff,
                                                       const int *__restrict__ dtot,
                                                       int *__restrict__ kout,
                                                       int *__restrict__ vout)
{
    __shared__ int run[256];
    __shared__ int wcnt[4][256];
    __shared__ int wsum[4];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int b = blockIdx.x;
    {
        int tot;
        const int dbase = block_exclusive_scan_256(dtot[t], wsum, &tot);
        run[t] = dbase + hoff[t * nblocks + b];
    }
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < kRadixBlock / 256; ++j) {
        const int i = b * kRadixBlock + j * 256 + t;
        const bool valid = i < n;
        const int k = valid ? kin[i] : 0;
        const int v = valid ? (vin ? vin[i] : i) : 0;
        const int d = (k >> shift) & 255;
        for (int w = 0; w < 4; ++w) wcnt[w][t] = 0;
        __syncthreads();
        // lanes of this wavefront holding the same digit
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1;
            const unsigned long long bal = __ballot(on);
            m &= on ? bal : ~bal;
        }
        const int rank = __popcll(m & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
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
    const int nb = radix_blocks(n);
    const int passes = (key_bits(nkeys) + 7) / 8;
    if (n > 0) {
        // ping-pong: key -> key2 -> key ...; values: identity -> tmp -> perm -> ...
        const int *kin = ks->key;
        const int *vin = nullptr;
        for (int p = 0; p < passes; ++p) {
            int *kout = (kin == ks->key) ? ks->key2 : ks->key;
            // values alternate so that the last pass lands in perm
            int *vout = ((passes - 1 - p) % 2 == 0) ? ks->perm : ks->tmp;
            hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, st, kin, (int)n, 8 * p, nb,
                               ks->hist);
            hipLaunchKernelGGL(k_radix_digit_scan, dim3(256), dim3(256), 0, st, ks->hist, nb,
                               ks->hoff, ks->hist + 256 * nb);
            hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, st, kin, vin, (int)n, 8 * p,
                               nb, ks->hoff, ks->hist + 256 * nb, kout, vout);
            kin = kout;
            vin = vout;
        }
        ks->sorted_key = const_cast<int *>(kin);
    } else {
        ks->sorted_key = ks->key;
    }
    hipLaunchKernelGGL(k_run_offsets, dim3((nkeys + 1 + 255) / 256), dim3(256), 0, st,
                       ks->sorted_key, (int)n, nkeys, ks->off);
    return 0;
}

// map[k] = c for k in [off[c], off[c+1]): workgroup index -> owning item.
// map[k] = c for every k in [off[c], off[c+1]), k < total = off[n]: one thread
// per entry, c by binary search over the n + 1 offsets (non-decreasing).
__global__ void k_expand_map(const int *__restrict__ off, int n, int total, int *__restrict__ map)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    int lo = 0, hi = n;   // off[lo] <= k < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= k) lo = mid;
        else hi = mid;
    }
    map[k] = lo;
}

void expand_map(hipStream_t st, const int *off, int n, int total, int *map)
{
    if (total > 0)
        hipLaunchKernelGGL(k_expand_map, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, off,
                           n, total, map);
}

// cell key (ewald_tools.cpp:28-35); with sublog > 0 the key also carries a
// 2^sublog x 2^sublog Morton-ordered sub-cell so that consecutive points of
// a cell are spatially close.
__global__ void k_cell_keys(const double *__restrict__ p, int n, double Lx, double Ly, int nsx,
                            int nsy, int sublog, int *__restrict__ key)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = se2p_cell_key(p[2 * i], p[2 * i + 1], Lx, Ly, nsx, nsy, sublog);
}

int cell_sort(se2p_ctx *ctx, const double *p, int64_t n, double Lx, double Ly, int nsx, int nsy,
              int sublog, KeySort *ks)
{
    hipStream_t st = ctx->stream;
    const int nkeys = nsx * nsy << (2 * sublog);
    if (n > 0)
        hipLaunchKernelGGL(k_cell_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p,
                           (int)n, Lx, Ly, nsx, nsy, sublog, ks->key);
    return keysort_run(ctx, ks, n, nkeys);
}

// Chunk plans (plan_body), one workgroup per plan (blockIdx.x selects plan 0
// or 1).  Replaces a count + scan + expand sequence with one launch.
__global__ __launch_bounds__(256) void k_plan(PlanArgs p0, PlanArgs p1)
{
    __shared__ int wsum[4];
    plan_body<4>(blockIdx.x == 0 ? p0 : p1, wsum);
}

void plan_chunks(hipStream_t st, const int *cnt, const int *roff, int sub, int per, int n, int *out,
                 int *map)
{
    PlanArgs p{cnt, roff, sub, per, n, out, map};
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, st, p, p);
}

void plan_chunks2(hipStream_t st, const int *cnt0, int n0, int *out0, int *map0, const int *cnt1,
                  int n1, int *out1, int *map1)
{
    PlanArgs p0{cnt0, nullptr, 1, 1, n0, out0, map0};
    PlanArgs p1{cnt1, nullptr, 1, 1, n1, out1, map1};
    hipLaunchKernelGGL(k_plan, dim3(2), dim3(256), 0, st, p0, p1);
}

}  // namespace se2p
