// se2p_fsort.h -- fused stable key sort for the common sizes (up to
// kFsMaxBlocks * kFsMaxRounds * 1024 points and kFsMaxKeys keys): four
// launches in place of the key / radix (3 per pass) / run-offset / gather
// launches of se2p_sort.hip, with the gather and the chunk planning riding
// along.
//
//   count    : the key of every point (KeyFn), a per-workgroup key histogram
//              in LDS, global key counts (integer atomics, one per key and
//              workgroup) and per-workgroup counts of the low 8-bit digit.
//   scan     : one workgroup: run offsets (off[0..nkeys]) from the key
//              counts, which it clears for the next sort; scatter offsets of
//              the low digit and the start of each low-digit group.
//   scatter 0: stable scatter by the low digit (fixed chunks of the input).
//   scatter 1: one workgroup per low-digit group d: the group holds every
//              key (h << 8) | d in input order, so a key's elements go to
//              off[key] onwards in the order met.  A payload functor writes
//              the sorted per-point data (positions, records).
// Within a workgroup, elements are ranked by ballots over the digit bits
// (equal digits of one wavefront), then by wavefront, then by round, so the
// order is the input order within each key: the same result as the serial
// counting sort of the cell lists, and the same on every run.
//
// Workgroups never wait for each other and need no device-scope fences:
// every result crosses to its readers at a kernel boundary.  The key counts
// live in a small per-context buffer that is zero between sorts (zeroed once
// when the context is created), so calls on one context must not overlap --
// as for its workspace.
//
// Extra workgroups appended to a scatter launch run work that needs only
// what the earlier launches produced (chunk plans, tile work lists).
#pragma once

#include "se2p_internal.h"

#include <algorithm>

namespace se2p {

constexpr int kFsThreads = 1024;   // threads per workgroup (16 wavefronts)
constexpr int kFsWaves = kFsThreads / 64;
// LDS stays under 64 KiB per workgroup: the key histogram (count launch) and
// the 256 x nb digit offsets (last workgroups) take at most 56 KiB.
constexpr int kFsMaxBlocks = 56;   // and kFsMaxKeys (se2p_internal.h)
constexpr int kFsMaxRounds = 4;    // rounds of 1024 elements per workgroup

// Whether the fused sort handles n points with nkeys keys on this context.
inline bool fsort_ok(const se2p_ctx *ctx, int64_t n, int nkeys)
{
    return ctx->fs != nullptr && !(ctx->sort_mode & SE2P_SORT_RADIX) && n > 0 &&
           n <= (int64_t)kFsMaxBlocks * kFsMaxRounds * kFsThreads && nkeys > 0 && nkeys <= kFsMaxKeys;
}

struct FsortArgs {
    int n, nkeys;
    int nb;        // workgroups of the count and scatter launches
    int rounds;    // rounds of kFsThreads elements per workgroup
    int *key;      // keys in input order (count)
    int *key2;     // keys after the low-digit pass
    int *tmp;      // input indices after the low-digit pass
    int *perm;     // sorted position -> input index
    int *off;      // [nkeys + 1] run offsets
    int *hist0;    // [nb][256] low-digit counts, then scatter offsets
    int *gstart;   // [257] start of each low-digit group after scatter 0
    int *cnt;      // [nkeys] key counts, zero between sorts
};

// Exclusive scan of one value per thread over a workgroup of NW wavefronts;
// *total receives the sum.
template <int NW>
__device__ inline int block_scan_excl(int v, int *wsum, int *total)
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
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int s = wsum[w];
        base += w < wave ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// Count launch.  Dynamic LDS: nkeys ints.
template <class KeyFn>
__global__ __launch_bounds__(kFsThreads) void k_fs_count(FsortArgs a, KeyFn kf)
{
    extern __shared__ int fs_lds[];
    int *lh = fs_lds;
    __shared__ int part[kFsThreads];
    const int t = threadIdx.x, b = blockIdx.x;
    for (int k = t; k < a.nkeys; k += kFsThreads) lh[k] = 0;
    __syncthreads();
    for (int j = 0; j < a.rounds; ++j) {
        const int i = (b * a.rounds + j) * kFsThreads + t;
        if (i < a.n) {
            // keys are in range by construction; the clamp keeps every LDS
            // and global access in bounds regardless
            const int k = min(max(kf(i), 0), a.nkeys - 1);
            a.key[i] = k;
            atomicAdd(&lh[k], 1);
        }
    }
    __syncthreads();
    // keys k = t (mod 1024) all have the low digit t & 255
    int d0 = 0;
    for (int k = t; k < a.nkeys; k += kFsThreads) {
        const int c = lh[k];
        if (c) {
            d0 += c;
            atomicAdd(&a.cnt[k], c);
        }
    }
    part[t] = d0;
    __syncthreads();
    if (t < 256) a.hist0[b * 256 + t] = part[t] + part[t + 256] + part[t + 512] + part[t + 768];
}

// Scan launch (one workgroup of 16 wavefronts): run offsets from the key
// counts (staged in LDS with coalesced loads; the counts are cleared for the
// next sort), the low digit's scatter offsets and group starts.  Thread t
// takes digit t & 255 of the workgroups in quarter t >> 8 (kept in
// registers).  Dynamic LDS: nkeys ints.
constexpr int kFsQuarter = (kFsMaxBlocks + 3) / 4;

template <int = 0>   // a template, so that every including file may instantiate it
__global__ __launch_bounds__(kFsThreads) void k_fs_scan(FsortArgs a)
{
    extern __shared__ int fs_lds[];
    int *kc = fs_lds;
    __shared__ int hp[4][256];
    __shared__ int dbase[256];
    __shared__ int wsum[kFsWaves];
    const int t = threadIdx.x;
    const int d = t & 255, q = t >> 8;
    const int nq = (a.nb + 3) >> 2;
    int hv[kFsQuarter];
    int ps = 0;
#pragma unroll
    for (int j = 0; j < kFsQuarter; ++j) {
        const int b = q * nq + j;
        hv[j] = (j < nq && b < a.nb) ? a.hist0[b * 256 + d] : 0;
        ps += hv[j];
    }
    for (int k = t; k < a.nkeys; k += kFsThreads) {
        kc[k] = a.cnt[k];
        a.cnt[k] = 0;
    }
    hp[q][d] = ps;
    __syncthreads();
    // run offsets: contiguous segments per thread, then a workgroup scan
    const int seg = (a.nkeys + kFsThreads - 1) / kFsThreads;
    const int k0 = min(t * seg, a.nkeys), k1 = min(k0 + seg, a.nkeys);
    int s = 0;
    for (int k = k0; k < k1; ++k) s += kc[k];
    int tot;
    int r = block_scan_excl<kFsWaves>(s, wsum, &tot);
    for (int k = k0; k < k1; ++k) {
        const int c = kc[k];
        kc[k] = r;
        r += c;
    }
    if (t == 0) a.off[a.nkeys] = tot;
    // digit starts
    const int dt = t < 256 ? hp[0][t] + hp[1][t] + hp[2][t] + hp[3][t] : 0;
    int all;
    const int db = block_scan_excl<kFsWaves>(dt, wsum, &all);
    if (t < 256) {
        dbase[t] = db;
        a.gstart[t] = db;
        if (t == 0) a.gstart[256] = all;
    }
    __syncthreads();
    int rr = dbase[d];
    for (int w = 0; w < q; ++w) rr += hp[w][d];
#pragma unroll
    for (int j = 0; j < kFsQuarter; ++j) {
        const int b = q * nq + j;
        if (j < nq && b < a.nb) a.hist0[b * 256 + d] = rr;
        rr += hv[j];
    }
    for (int k = t; k < a.nkeys; k += kFsThreads) a.off[k] = kc[k];
}

// Low-digit scatter: workgroup b takes input chunk b.  Workgroups nb.. run
// the extra work.  Dynamic LDS: kFsWaves*256 ints (per-wavefront digit
// counts).
template <class Extra>
__global__ __launch_bounds__(kFsThreads) void k_fs_scatter0(FsortArgs a, Extra ex)
{
    extern __shared__ int fs_lds[];
    int(*wcnt)[256] = (int(*)[256])fs_lds;
    __shared__ int wsum[kFsWaves];
    const int b = blockIdx.x;
    if (b >= a.nb) {   // uniform over the workgroup
        ex.template run<kFsWaves>(b - a.nb, wsum);
        return;
    }
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int run = t < 256 ? a.hist0[b * 256 + t] : 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < a.rounds; ++j) {
        const int i = (b * a.rounds + j) * kFsThreads + t;
        const bool valid = i < a.n;
        const int k = valid ? a.key[i] : 0;
        const int d = k & 255;
        for (int x = t; x < kFsWaves * 256; x += kFsThreads) fs_lds[x] = 0;
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
        if (t < 256) {   // per digit: start of each wavefront's elements
#pragma unroll
            for (int w = 0; w < kFsWaves; ++w) {
                const int c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
        }
        __syncthreads();
        if (valid) {
            const int pos = wcnt[wave][d] + rank;
            a.key2[pos] = k;
            a.tmp[pos] = i;
        }
        __syncthreads();
    }
}

// Final scatter: workgroup d takes low-digit group d (gstart[d] ..
// gstart[d+1]) in rounds of 1024; the element of key k = (h << 8) | d goes
// to off[k] + (its rank among the group's earlier elements of key k).
// Workgroups 256.. run the extra work.  (1024 threads finish a group of a
// few hundred points in one round: 31-34 against 39 us for the benchmark's
// cell-list sort, 27 against 30-32 us for its bin sort, with 256.)
constexpr int kFsFinalThreads = 1024;
constexpr int kFsMaxHigh = (kFsMaxKeys + 255) / 256;   // high digits (<= 64)
static_assert(kFsMaxHigh <= 64, "six ballots rank the high digit");

// Order-preserving map of a double (not NaN) to a signed 64-bit integer,
// and back: integer min / max then equal the floating-point ones.
__device__ inline long long fs_ord(double x)
{
    const long long b = __double_as_longlong(x);
    return b >= 0 ? b : b ^ 0x7fffffffffffffffLL;
}
__device__ inline double fs_unord(long long m)
{
    return __longlong_as_double(m >= 0 ? m : m ^ 0x7fffffffffffffffLL);
}

// With Payload::kBox, the payload returns the point's position and the
// final scatter writes every key's bounding box (xmin, ymin, xmax, ymax) to
// pay.box[key] -- (+inf, +inf, -inf, -inf) for an empty key.  A workgroup
// holds all keys of its low digit, so the boxes are complete in its LDS.
template <class Payload, class Extra>
__global__ __launch_bounds__(kFsFinalThreads) void k_fs_final(FsortArgs a, Payload pay, Extra ex)
{
    constexpr int NW = kFsFinalThreads / 64;
    constexpr bool BOX = Payload::kBox;
    __shared__ int wcnt[NW][64];
    __shared__ int wsum[NW];
    __shared__ long long bb[BOX ? 4 * kFsMaxHigh : 1];
    const int d = blockIdx.x;
    if (d >= 256) {   // uniform over the workgroup
        ex.template run<NW>(d - 256, wsum);
        return;
    }
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int g0 = a.gstart[d], g1 = a.gstart[d + 1];
    const int nhi = (a.nkeys + 255) >> 8;
    if (BOX) {
        if (t < nhi) {
            const double inf = __builtin_inf();
            bb[4 * t] = fs_ord(inf);
            bb[4 * t + 1] = fs_ord(inf);
            bb[4 * t + 2] = fs_ord(-inf);
            bb[4 * t + 3] = fs_ord(-inf);
        }
    } else if (g0 >= g1) {
        return;   // empty group (uniform)
    }
    int run = 0;
    if (t < nhi) {
        const int k = (t << 8) | d;
        run = k < a.nkeys ? a.off[k] : 0;   // keys past nkeys never occur
    }
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int base = g0; base < g1; base += kFsFinalThreads) {
        const int p0 = base + t;
        const bool valid = p0 < g1;
        const int k = valid ? a.key2[p0] : 0;
        const int v = valid ? a.tmp[p0] : 0;
        const int h = k >> 8;
        wcnt[wave][lane] = 0;
        __syncthreads();
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 6; ++bit) {
            const bool on = (h >> bit) & 1;
            const unsigned long long bal = __ballot(on);
            m &= on ? bal : ~bal;
        }
        const int rank = __popcll(m & lt);
        if (valid && rank == 0) wcnt[wave][h] = __popcll(m);
        __syncthreads();
        if (t < nhi) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const int c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
        }
        __syncthreads();
        if (valid) {
            const int pos = wcnt[wave][h] + rank;
            a.perm[pos] = v;
            if constexpr (BOX) {
                const double2 p = pay(pos, v, k);
                atomicMin(&bb[4 * h], fs_ord(p.x));
                atomicMin(&bb[4 * h + 1], fs_ord(p.y));
                atomicMax(&bb[4 * h + 2], fs_ord(p.x));
                atomicMax(&bb[4 * h + 3], fs_ord(p.y));
            } else {
                pay(pos, v, k);
            }
        }
        __syncthreads();
    }
    if constexpr (BOX) {
        if (t < nhi) {
            const int k = (t << 8) | d;
            if (k < a.nkeys)
                pay.box[k] = make_double4(fs_unord(bb[4 * t]), fs_unord(bb[4 * t + 1]),
                                          fs_unord(bb[4 * t + 2]), fs_unord(bb[4 * t + 3]));
        }
    }
}

// Chunk plan: out[0..n] = exclusive scan of the counts and map[k] = c for
// every k in [out[c], out[c+1]).  The count of item c is cnt[c], or with
// cnt == null ceil((roff[(c+1)*sub] - roff[c*sub]) / per), the chunks of
// `per` points of a run.  One workgroup of NW wavefronts.
struct PlanArgs {
    const int *cnt, *roff;
    int sub, per, n;
    int *out, *map;
};

template <int NW>
__device__ inline void plan_body(const PlanArgs &p, int *wsum)
{
    int carry = 0;
    for (int base = 0; base < p.n; base += 64 * NW) {
        const int c = base + threadIdx.x;
        int v = 0;
        if (c < p.n)
            v = p.cnt ? p.cnt[c] : (p.roff[(c + 1) * p.sub] - p.roff[c * p.sub] + p.per - 1) / p.per;
        int tot;
        const int ex = block_scan_excl<NW>(v, wsum, &tot);
        if (c < p.n) {
            p.out[c] = carry + ex;
            for (int k = 0; k < v; ++k) p.map[carry + ex + k] = c;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) p.out[p.n] = carry;
}

// Extra work: one or two chunk plans (one workgroup each).
struct FsPlanExtra {
    PlanArgs p0, p1;
    int nplans;   // 0, 1 or 2
    int blocks(int) const { return nplans; }
    template <int NW>
    __device__ void run(int xb, int *wsum) const
    {
        plan_body<NW>(xb == 0 ? p0 : p1, wsum);
    }
};

struct FsNoPayload {
    static constexpr bool kBox = false;
    __device__ void operator()(int, int, int) const {}
};
struct FsNoExtra {
    int blocks(int) const { return 0; }
    template <int NW>
    __device__ void run(int, int *) const {}
};

// Sort n points by kf's keys: ks->perm and ks->off filled, the payload
// called once per point with (sorted position, input index, key).  Extra
// work x0 runs beside the low-digit scatter (run offsets ready), x1 beside
// the final scatter (x0's results ready).
template <class KeyFn, class Payload, class Extra0, class Extra1>
int fsort_run(se2p_ctx *ctx, KeySort *ks, int64_t n, int nkeys, const KeyFn &kf,
              const Payload &pay, const Extra0 &x0, const Extra1 &x1)
{
    hipStream_t st = ctx->stream;
    FsortArgs a;
    a.n = (int)n;
    a.nkeys = nkeys;
    a.rounds = (int)((n + (int64_t)kFsMaxBlocks * kFsThreads - 1) / ((int64_t)kFsMaxBlocks * kFsThreads));
    a.nb = (int)((n + (int64_t)a.rounds * kFsThreads - 1) / ((int64_t)a.rounds * kFsThreads));
    a.key = ks->key;
    a.key2 = ks->key2;
    a.tmp = ks->tmp;
    a.perm = ks->perm;
    a.off = ks->off;
    a.hist0 = ks->hist;
    a.gstart = ks->hoff;
    a.cnt = ctx->fs;
    const size_t lds_count = sizeof(int) * (size_t)nkeys;
    const size_t lds_scan = sizeof(int) * (size_t)nkeys;
    const size_t lds_scatter = sizeof(int) * 256 * (size_t)kFsWaves;
    hipLaunchKernelGGL((k_fs_count<KeyFn>), dim3(a.nb), dim3(kFsThreads), lds_count, st, a, kf);
    hipLaunchKernelGGL(k_fs_scan<>, dim3(1), dim3(kFsThreads), lds_scan, st, a);
    hipLaunchKernelGGL((k_fs_scatter0<Extra0>), dim3(a.nb + x0.blocks(kFsWaves)), dim3(kFsThreads),
                       lds_scatter, st, a, x0);
    hipLaunchKernelGGL((k_fs_final<Payload, Extra1>), dim3(256 + x1.blocks(kFsFinalThreads / 64)),
                       dim3(kFsFinalThreads), 0, st, a, pay, x1);
    ks->sorted_key = nullptr;   // not produced
    SE2P_HIP(hipGetLastError());
    return 0;
}

}  // namespace se2p
