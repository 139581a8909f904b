// se2p_internal.h -- context, workspace and launch helpers shared by the
// spectral Ewald HIP sources.  Not part of the public C ABI (include/se2p.h).
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/se2p.h"

namespace se2p {

// Number of timed stages recorded when timing is enabled.
enum Stage {
    ST_SORT = 0,      // cell lists / gridding bins
    ST_REAL,          // real-space pair sum
    ST_SPREAD,        // Gaussian spreading
    ST_FFT_FWD,       // forward transforms
    ST_FILTER,        // k-space filter
    ST_FFT_INV,       // inverse transforms
    ST_INTERP,        // Gaussian interpolation
    ST_PAIR,          // the real-space pair kernel alone (inside ST_REAL)
    ST_COUNT
};

struct FftPlans {
    rocfft_plan fwd = nullptr;
    rocfft_plan inv = nullptr;
    rocfft_execution_info fwd_info = nullptr;
    rocfft_execution_info inv_info = nullptr;
    size_t fwd_work = 0, inv_work = 0;
    void *work = nullptr;   // shared work buffer (hipMalloc, freed with the context)
};

// State a forward call keeps for its backward call (se2p_set_keep); defined
// in se2p_real.hip and se2p_kspace.hip.
struct KeptReal;
struct KeptK;

}  // namespace se2p

struct se2p_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool async = false;
    bool timing = false;
    char *ws = nullptr;       // device workspace
    size_t ws_bytes = 0;
    std::string err;
    std::map<std::tuple<int, int, int, int, int>, se2p::FftPlans> plans;  // (Mx, My, nfwd, ninv, slot)
    // stage timing: up to kStageVisits (start, end) event pairs per stage and call
    static constexpr int kStageVisits = 4;
    hipEvent_t ev[se2p::ST_COUNT][kStageVisits][2] = {};
    int ev_used[se2p::ST_COUNT] = {};   // visits recorded in this call
    float stage_ms[se2p::ST_COUNT] = {};
    int64_t pairs_in = -1, pairs_out = -1;   // counted by se2p_count_pairs
    char *drv = nullptr;      // driver-level buffers (kept apart from ws)
    size_t drv_bytes = 0;
    bool host_ptrs = false;   // array arguments are host memory (se2p_set_host_pointers)
    char *io = nullptr;       // device staging area of host-pointer calls
    size_t io_bytes = 0;
    int *fs = nullptr;        // fused-sort state (se2p_fsort.h), zero between sorts
    int sort_mode = 0;        // SE2P_SORT_* bits
    int spread_chunk = 0;     // candidate sources per spreading workgroup (0: the default)
    // Stream slots for independent work run side by side (multi-box Fourier
    // sums, concurrent groups): each has its own stream, workspace, sort
    // state and FFT plans
    // (plan_slot in the plan key).  Slot work forks from and joins into
    // `stream` through events.
    struct Slot {
        hipStream_t st = nullptr;
        char *ws = nullptr;
        size_t ws_bytes = 0;
        int *fs = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<Slot> slots;    // multi-box Fourier sums
    std::vector<Slot> gslots;   // concurrent groups and driver calls
    hipEvent_t fork_ev = nullptr;
    hipEvent_t switch_ev = nullptr;   // orders a newly bound stream after the previous one
    // se2p_group_begin .. se2p_group_end: -1 outside a group, else the number
    // of calls made in the open group; call 0 runs on `stream`, call k >= 1 on
    // gslots[k - 1], which waits for group_ev (recorded on `stream` at
    // se2p_group_begin)
    int group = -1;
    hipEvent_t group_ev = nullptr;
    bool beside = false;      // a driver call runs its two sums side by side
    int plan_slot = 0;        // 0: the context's own plans
    bool keep = false;        // forward calls keep their sorted state (se2p_set_keep)
    bool split_adjoint = false;   // se2p_set_adjoint_passes(ctx, 2)
    se2p::KeptReal *kept_real = nullptr;
    se2p::KeptK *kept_k = nullptr;
    // This device's address of the fixed scratch of se2p_dlp_real_noalloc
    // (looked up once per device at se2p_create; null if that failed).
    char *scratch = nullptr;
    int ncu = 256;            // compute units of the device (persistent launches)
    // se2p_set_frozen: calls that would allocate or free device memory fail
    // with SE2P_ERR_FROZEN instead (captured graphs keep their addresses)
    bool frozen = false;
};

namespace se2p {

// Fused-sort state per context: the key counts of up to kFsMaxKeys keys.
constexpr int kFsMaxKeys = 14336;

int fail(se2p_ctx *ctx, int code, const char *fmt, ...);
// 0, or SE2P_ERR_FROZEN (with a message naming `what`) on a frozen context.
int refuse_if_frozen(se2p_ctx *ctx, const char *what, size_t bytes);

// Ensure the workspace holds at least `bytes`.
int ensure_workspace(se2p_ctx *ctx, size_t bytes);

// Stream slots: at least n of them, each with its stream, event and sort
// state.  Multi-box Fourier-space sums use kKspaceSlots of ctx->slots;
// concurrent groups and the driver entries use ctx->gslots (`group`), so
// that neither overwrites the other's workspace or plans.
int ensure_slots(se2p_ctx *ctx, int n, bool group = false);
constexpr int kKspaceSlots = 4;

inline se2p_ctx::Slot &slot_at(se2p_ctx *ctx, int k, bool group = false)
{
    return group ? ctx->gslots[(size_t)k] : ctx->slots[(size_t)k];
}

// The context's stream, workspace, sort state and plan set swapped for slot
// k's while in scope (restored on every exit path).
struct SlotScope {
    se2p_ctx *ctx;
    se2p_ctx::Slot &slot;
    hipStream_t st;
    char *ws;
    size_t ws_bytes;
    int *fs;
    bool timing;
    SlotScope(se2p_ctx *c, int k, bool group = false)
        : ctx(c), slot(slot_at(c, k, group)), st(c->stream), ws(c->ws), ws_bytes(c->ws_bytes),
          fs(c->fs), timing(c->timing)
    {
        ctx->stream = slot.st;
        ctx->ws = slot.ws;
        ctx->ws_bytes = slot.ws_bytes;
        ctx->fs = slot.fs;
        ctx->timing = false;
        ctx->plan_slot = group ? kKspaceSlots + 1 + k : k + 1;
    }
    ~SlotScope()
    {
        slot.ws = ctx->ws;   // ensure_workspace may have grown it
        slot.ws_bytes = ctx->ws_bytes;
        ctx->stream = st;
        ctx->ws = ws;
        ctx->ws_bytes = ws_bytes;
        ctx->fs = fs;
        ctx->timing = timing;
        ctx->plan_slot = 0;
    }
};

// Whether other work of the caller runs beside this call on the GPU (an open
// concurrent group, or the driver entries' two sums): persistent kernels then
// leave part of each CU to it.
inline bool shares_gpu(const se2p_ctx *ctx) { return ctx->group >= 0 || ctx->beside; }

// Timing helpers (no-ops unless ctx->timing).
void stage_begin(se2p_ctx *ctx, Stage s);
void stage_end(se2p_ctx *ctx, Stage s);

int finish(se2p_ctx *ctx);   // error check + optional synchronisation

// Release the kept forward state (se2p_real.hip, se2p_kspace.hip).
void free_kept_real(se2p_ctx *ctx);
void free_kept_k(se2p_ctx *ctx);

// Bump allocator over the workspace, 256-byte aligned pieces.  With a null
// base it only measures: run the same sequence of take() calls once to size
// the workspace and once to hand out pointers.
struct Carve {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    T *take(size_t count)
    {
        T *p = base ? (T *)(base + off) : nullptr;
        off += (sizeof(T) * (count > 0 ? count : 1) + 255) & ~size_t(255);
        return p;
    }
};

// Zero n doubles with a fill kernel (no memset node, so capturable).
void launch_zero(hipStream_t st, double *p, int64_t n);

// ---- cell lists / bins (se2p_sort.hip) ---------------------------------
// Stable sort of points by an integer key in [0, nkeys): perm[k] = original
// index of the k-th point, off[0..nkeys] = start of each key's run.
struct KeySort {
    int *key;          // keys (input; also a ping-pong buffer)
    int *key2, *tmp;   // ping-pong buffers
    int *perm;         // sorted order -> original index
    int *off;          // [nkeys+1] run offsets
    int *hist, *hoff;  // per-block digit histograms and their scan
    int *sorted_key;   // the buffer holding the sorted keys after the sort
};
size_t keysort_bytes(int64_t n, int nkeys);
void keysort_carve(char *base, int64_t n, int nkeys, KeySort *ks);
// Sort ks->key[0..n) (filled by the caller), producing perm and off.
int keysort_run(se2p_ctx *ctx, KeySort *ks, int64_t n, int nkeys);
// Cell lists (ewald_tools.cpp:28-35); sublog > 0 appends a Morton sub-cell of
// 2^sublog x 2^sublog to the key (nkeys = nsx*nsy*4^sublog).
int cell_sort(se2p_ctx *ctx, const double *p, int64_t n, double Lx, double Ly, int nsx, int nsy,
              int sublog, KeySort *ks);
// Exclusive scan of n ints into out[0..n] with one workgroup.
void scan_ints(hipStream_t st, const int *a, int n, int *out);
// map[k] = c for every k in [off[c], off[c+1]), c < n; total = off[n].
void expand_map(hipStream_t st, const int *off, int n, int total, int *map);
// One launch: out[0..n] = exclusive scan of the counts, map[k] = owner c of
// chunk k.  Counts are cnt[c], or (cnt null) ceil(run length / per) of the
// runs roff[c*sub] .. roff[(c+1)*sub].
void plan_chunks(hipStream_t st, const int *cnt, const int *roff, int sub, int per, int n, int *out,
                 int *map);
// Two independent count arrays planned in one launch.
void plan_chunks2(hipStream_t st, const int *cnt0, int n0, int *out0, int *map0, const int *cnt1,
                  int n1, int *out1, int *map1);

// ---- real space (se2p_real.hip) -----------------------------------------
struct RealGeom {
    const double *psrc, *ptar;
    int64_t ns, nt;
    double xi, Lx, Ly;
    int nsx, nsy;
};
int real_sum(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
             int nrhs, int64_t dstride, int64_t ustride, double *u, int64_t *pair_counts);
// Real-space sum over several independent boxes (host records of the boxes;
// se2p_real.hip fills in their global bases).
struct RealBox;
// fixed != null: the sorted state is carved from fixed[0..fixed_bytes) and
// nothing is allocated (SE2P_ERR_ARG if it does not fit).
int real_sum_boxes(se2p_ctx *ctx, int kind, RealBox *boxes, int nbox, int nrhs, int64_t dstride,
                   int64_t ustride, int64_t *pair_counts, char *fixed = nullptr,
                   size_t fixed_bytes = 0);
int real_sum_multibox(se2p_ctx *ctx, int kind, const se2p_box_desc *boxes, int nbox);
// The no-allocation form of real_sum (one density): the bytes of sorted state
// a call needs, and the call itself with that state in caller-given scratch.
int real_scratch_bytes(se2p_ctx *ctx, int kind, const RealGeom &g, size_t *bytes);
int real_sum_fixed(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
                   double *u, char *scratch_base, size_t scratch_bytes);
// Looks up this device's address of the fixed scratch into ctx->scratch
// (se2p_real.hip, where the array lives); called by se2p_create.
int noalloc_scratch_init(se2p_ctx *ctx);
// Adjoint of real_sum (one right-hand side): with L = sum_t du_t . u_t, the
// gradients dL/dpsrc, dL/dptar, dL/df and (stresslet) dL/dn; null outputs are
// skipped.
int real_adjoint(se2p_ctx *ctx, int kind, const RealGeom &g, const double *f, const double *n,
                 const double *du, double *dpsrc, double *dptar, double *df, double *dn);

// ---- Fourier space (se2p_kspace.hip) ------------------------------------
struct KGeom {
    const double *psrc, *ptar;
    int64_t ns, nt;
    double xi, eta, Lx, Ly, w;
    int Mx, My, P;
};
int kspace_sum(se2p_ctx *ctx, int kind, const KGeom &g, const double *f, const double *n,
               int nrhs, int64_t dstride, int64_t ustride, double *u);
// Fourier-space sums of many independent boxes (one right-hand side each).
int kspace_sum_boxes(se2p_ctx *ctx, int kind, const se2p_box_desc *boxes, int nbox);
// Adjoint of kspace_sum (one right-hand side), outputs as for real_adjoint.
int kspace_adjoint(se2p_ctx *ctx, int kind, const KGeom &g, const double *f, const double *n,
                   const double *du, double *dpsrc, double *dptar, double *df, double *dn);

// Validation of the k-space preconditions (returns 0 or an SE2P_ERR code).
int check_kspace_params(se2p_ctx *ctx, const KGeom &g);

// ---- driver entry points (se2p_driver.cpp, se2p_compose.hip) ------------
int select_params(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                  se2p_params *p);
void launch_wrap(hipStream_t st, const double *p, int64_t n, double Lx, double Ly, double *out);
// out[0] = sum f.f, out[1..2] = sum (f.n) x (nv may be null: then 0)
void launch_driver_sums(hipStream_t st, const double *f, const double *nv, const double *x,
                        int64_t n, double *partial, double *out);
void launch_compose(hipStream_t st, const double *ur, double *uk, int64_t nt, int dlp,
                    const double *zm, double LxLy, double *u);

}  // namespace se2p

#define SE2P_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return se2p::fail(ctx, SE2P_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                              hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)
