// se2p_api.cpp -- the C ABI (include/se2p.h): contexts, argument validation
// and the four MEX-equivalent entry points.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>
#include <mutex>
#include <optional>

#include "se2p_device.h"
#include "se2p_internal.h"

#include <limits>

namespace se2p {

int fail(se2p_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int refuse_if_frozen(se2p_ctx *ctx, const char *what, size_t bytes)
{
    if (!ctx->frozen) return 0;
    return fail(ctx, SE2P_ERR_FROZEN,
                "context is frozen (se2p_set_frozen): %s would allocate or free device memory "
                "(%zu bytes); make the same call once before freezing so that it is sized",
                what, bytes);
}

int ensure_workspace(se2p_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->ws_bytes) return 0;
    if (int rc = refuse_if_frozen(ctx, "growing the workspace", bytes)) return rc;
    if (ctx->ws) {
        SE2P_HIP(hipStreamSynchronize(ctx->stream));
        SE2P_HIP(hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    size_t grow = bytes + bytes / 4;
    SE2P_HIP(hipMalloc((void **)&ctx->ws, grow));
    ctx->ws_bytes = grow;
    return 0;
}

// A stage may run several times in one call (the adjoint's two passes); each
// visit gets its own event pair and the reported time is the sum of the visits.
void stage_begin(se2p_ctx *ctx, Stage s)
{
    if (!ctx->timing || ctx->ev_used[s] >= se2p_ctx::kStageVisits) return;
    hipEventRecord(ctx->ev[s][ctx->ev_used[s]][0], ctx->stream);
}

void stage_end(se2p_ctx *ctx, Stage s)
{
    if (!ctx->timing || ctx->ev_used[s] >= se2p_ctx::kStageVisits) return;
    hipEventRecord(ctx->ev[s][ctx->ev_used[s]][1], ctx->stream);
    ++ctx->ev_used[s];
}

int finish(se2p_ctx *ctx)
{
    SE2P_HIP(hipGetLastError());
    // a grouped call only enqueues; se2p_group_end joins and waits
    if (ctx->group >= 0) return SE2P_OK;
    if (!ctx->async || ctx->timing) SE2P_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->timing) {
        for (int s = 0; s < ST_COUNT; ++s) {
            ctx->stage_ms[s] = -1.0f;
            if (ctx->ev_used[s] > 0) {
                float total = 0.0f;
                for (int v = 0; v < ctx->ev_used[s]; ++v) {
                    float ms = 0.0f;
                    hipEventElapsedTime(&ms, ctx->ev[s][v][0], ctx->ev[s][v][1]);
                    total += ms;
                }
                ctx->stage_ms[s] = total;
            }
        }
    }
    return SE2P_OK;
}

static void begin_call(se2p_ctx *ctx)
{
    ctx->err.clear();
    for (int s = 0; s < ST_COUNT; ++s) ctx->ev_used[s] = 0;
}

static int set_device(se2p_ctx *ctx)
{
    if (ctx->device == SE2P_HOST_ONLY)
        return fail(ctx, SE2P_ERR_HIP, "host-only context: no device to run on");
    SE2P_HIP(hipSetDevice(ctx->device));
    return 0;
}

static int check_common(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                        int64_t nt, const double *u)
{
    if (ns < 0 || nt < 0) return fail(ctx, SE2P_ERR_ARG, "negative point count");
    if (ns > 0x7fffffffLL || nt > 0x7fffffffLL)
        return fail(ctx, SE2P_ERR_ARG, "point counts above 2^31-1 are not supported");
    if ((ns > 0 && !psrc) || (nt > 0 && (!ptar || !u)))
        return fail(ctx, SE2P_ERR_ARG, "null array pointer");
    return 0;
}

static int check_real(se2p_ctx *ctx, double xi, int nsx, int nsy, double Lx, double Ly)
{
    if (nsx < 1 || nsy < 1) return fail(ctx, SE2P_ERR_NSIDE, "nside_x and nside_y must be >= 1");
    // cell-list keys are (cell, 4x4 sub-cell) pairs in 32-bit ints: every
    // entry (forward, batched, backward, pair count) needs nsx*nsy*16 to fit
    if ((long long)nsx * nsy * 16 > 0x7fffffffLL)
        return fail(ctx, SE2P_ERR_NSIDE,
                    "too many cells: nside_x * nside_y * 16 must fit in a 32-bit int");
    if (!(xi > 0) || !(Lx > 0) || !(Ly > 0) || !std::isfinite(xi) || !std::isfinite(Lx) ||
        !std::isfinite(Ly))
        return fail(ctx, SE2P_ERR_PARAM, "xi, Lx and Ly must be positive and finite");
    return 0;
}

// Backward entries: du is required, every output may be null (skipped).
static int check_adjoint(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                         int64_t nt, const double *f, const double *n, bool dlp, const double *du)
{
    if (ns < 0 || nt < 0) return fail(ctx, SE2P_ERR_ARG, "negative point count");
    if (ns > 0x7fffffffLL || nt > 0x7fffffffLL)
        return fail(ctx, SE2P_ERR_ARG, "point counts above 2^31-1 are not supported");
    if ((ns > 0 && (!psrc || !f || (dlp && !n))) || (nt > 0 && (!ptar || !du)))
        return fail(ctx, SE2P_ERR_ARG, "null array pointer");
    return 0;
}

// Host-pointer mode (se2p_set_host_pointers): the call's arrays are host
// memory.  Inputs are copied into a device staging area before the
// computation and outputs back after it; such a call is always synchronous.
// A host array passed twice (targets that are the sources) maps to one device
// copy, so the shared cell lists still apply.  In device mode every method
// returns its argument unchanged.
class HostIo {
public:
    explicit HostIo(se2p_ctx *ctx) : ctx_(ctx), on_(ctx->host_ptrs) {}

    // declare an array (input or output) and its size in bytes; an array
    // declared twice (ptar == psrc) is staged once, at the larger size
    void want(const void *p, size_t bytes)
    {
        if (!on_ || !p || bytes == 0) return;
        if (Item *it = find(p)) {
            if (bytes > it->bytes) {
                need_ += ((bytes + 255) & ~size_t(255)) - ((it->bytes + 255) & ~size_t(255));
                it->bytes = bytes;
            }
            return;
        }
        items_.push_back({p, nullptr, bytes, false});
        need_ += (bytes + 255) & ~size_t(255);
    }
    // allocate the staging area for everything declared
    int map()
    {
        if (!on_) return 0;
        if (need_ > ctx_->io_bytes) {
            if (int rc = refuse_if_frozen(ctx_, "growing the host-pointer staging area", need_))
                return rc;
            if (ctx_->io) {
                SE2P_HIP(hipStreamSynchronize(ctx_->stream));
                SE2P_HIP(hipFree(ctx_->io));
                ctx_->io = nullptr;
                ctx_->io_bytes = 0;
            }
            SE2P_HIP(hipMalloc((void **)&ctx_->io, need_));
            ctx_->io_bytes = need_;
        }
        size_t o = 0;
        for (Item &it : items_) {
            it.dev = ctx_->io + o;
            o += (it.bytes + 255) & ~size_t(255);
        }
        return 0;
    }
    template <class T>
    const T *in(const T *p)
    {
        Item *it = on_ && p ? find(p) : nullptr;
        if (!it) return p;
        if (!it->copied) {
            hipError_t e = hipMemcpyAsync(it->dev, p, it->bytes, hipMemcpyHostToDevice, ctx_->stream);
            if (e != hipSuccess) err_ = e;
            it->copied = true;
        }
        return (const T *)it->dev;
    }
    template <class T>
    T *out(T *p)
    {
        Item *it = on_ && p ? find(p) : nullptr;
        if (!it) return p;
        outs_.push_back(it);
        return (T *)it->dev;
    }
    // after the computation: copy the outputs back and wait for them
    int done(int rc)
    {
        if (!on_ || rc != SE2P_OK) return rc;
        SE2P_HIP(err_);
        for (Item *it : outs_)
            SE2P_HIP(hipMemcpyAsync((void *)it->host, it->dev, it->bytes, hipMemcpyDeviceToHost,
                                    ctx_->stream));
        SE2P_HIP(hipStreamSynchronize(ctx_->stream));
        return SE2P_OK;
    }

private:
    struct Item {
        const void *host;
        char *dev;
        size_t bytes;
        bool copied;
    };
    Item *find(const void *p)
    {
        for (Item &it : items_)
            if (it.host == p) return &it;
        return nullptr;
    }
    se2p_ctx *ctx_;
    bool on_;
    size_t need_ = 0;
    hipError_t err_ = hipSuccess;
    std::vector<Item> items_;
    std::vector<Item *> outs_;
    se2p_ctx *ctx = ctx_;   // for SE2P_HIP
};

int ensure_slots(se2p_ctx *ctx, int n, bool group)
{
    std::vector<se2p_ctx::Slot> &pool = group ? ctx->gslots : ctx->slots;
    if ((int)pool.size() < n)
        if (int rc = refuse_if_frozen(ctx, "creating stream slots", sizeof(int) * kFsMaxKeys)) return rc;
    if (!ctx->fork_ev) SE2P_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    while ((int)pool.size() < n) {
        se2p_ctx::Slot sl;
        SE2P_HIP(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        SE2P_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        const size_t fs_bytes = sizeof(int) * kFsMaxKeys;
        SE2P_HIP(hipMalloc((void **)&sl.fs, fs_bytes));
        SE2P_HIP(hipMemsetAsync(sl.fs, 0, fs_bytes, sl.st));
        SE2P_HIP(hipStreamSynchronize(sl.st));
        pool.push_back(sl);
    }
    return 0;
}

// SE2P_ERR_ARG while a group (se2p_group_begin) is open, naming `what`.
static int refuse_in_group(se2p_ctx *ctx, const char *what)
{
    if (ctx->group < 0) return 0;
    return fail(ctx, SE2P_ERR_ARG, "%s is not allowed between se2p_group_begin and se2p_group_end",
                what);
}

// A call made while a group is open: the group's first call runs on the
// context stream itself (which is where group_ev was recorded, so it needs
// no wait), call k >= 1 on slot k - 1, whose stream waits for group_ev and
// whose stream, workspace, sort state and plan set are the context's until
// the call returns; the slot's completion event is recorded on every exit
// path (se2p_group_end waits for it).  Stage timing is off for grouped
// calls.  Outside a group enter() does nothing.
class GroupCall {
public:
    explicit GroupCall(se2p_ctx *c) : ctx(c) {}
    int enter()
    {
        if (ctx->group < 0) return 0;
        if (ctx->host_ptrs)
            return fail(ctx, SE2P_ERR_ARG,
                        "host-pointer calls are synchronous and cannot run in a group");
        if (ctx->group >= SE2P_GROUP_MAX_CALLS)
            return fail(ctx, SE2P_ERR_ARG, "a group holds at most %d calls", SE2P_GROUP_MAX_CALLS);
        const int k = ctx->group;
        if (k > 0) {
            if (int rc = ensure_slots(ctx, k, true)) return rc;
            SE2P_HIP(hipStreamWaitEvent(ctx->gslots[(size_t)k - 1].st, ctx->group_ev, 0));
        }
        ctx->group = k + 1;
        k_ = k;
        if (ctx->timing)
            for (float &ms : ctx->stage_ms) ms = -1.0f;
        timing_ = ctx->timing;
        if (k > 0) scope_.emplace(ctx, k - 1, true);   // it turns timing off itself
        else ctx->timing = false;
        return 0;
    }
    ~GroupCall()
    {
        if (k_ < 0) return;
        if (k_ == 0) {
            ctx->timing = timing_;
            return;
        }
        scope_.reset();
        hipEventRecord(ctx->gslots[(size_t)k_ - 1].done, ctx->gslots[(size_t)k_ - 1].st);
    }

private:
    se2p_ctx *ctx;
    int k_ = -1;
    bool timing_ = false;
    std::optional<SlotScope> scope_;
};

// The Fourier-space sum `fourier` on the context stream and the real-space
// sum `real` beside it on group slot 0, as the two calls of a concurrent
// group (Fourier-space call first) would run; joined into the context stream
// also after an error, so that the stream stays ordered.
template <class F, class R>
static int sums_side_by_side(se2p_ctx *ctx, F &&fourier, R &&real)
{
    hipStream_t st = ctx->stream;
    if (int rc = ensure_slots(ctx, 1, true)) return rc;
    se2p_ctx::Slot &sl = ctx->gslots[0];
    SE2P_HIP(hipEventRecord(ctx->fork_ev, st));
    SE2P_HIP(hipStreamWaitEvent(sl.st, ctx->fork_ev, 0));
    const bool beside = ctx->beside;
    ctx->beside = true;
    int rc = fourier();
    if (rc == 0) {
        SlotScope scope(ctx, 0, true);
        rc = real();
    }
    ctx->beside = beside;
    SE2P_HIP(hipEventRecord(sl.done, sl.st));
    SE2P_HIP(hipStreamWaitEvent(st, sl.done, 0));
    return rc;
}

}  // namespace se2p

using namespace se2p;

extern "C" {

int se2p_create(se2p_ctx **out, int device)
{
    if (!out) return SE2P_ERR_ARG;
    *out = nullptr;
    static std::once_flag fft_once;
    std::call_once(fft_once, [] { rocfft_setup(); });
    se2p_ctx *ctx = new se2p_ctx();
    ctx->device = device;
    if (device == SE2P_HOST_ONLY) {
        // argument validation only; every compute call fails with SE2P_ERR_HIP
        *out = ctx;
        return SE2P_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        delete ctx;
        return SE2P_ERR_HIP;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return SE2P_ERR_HIP;
    }
    ctx->own_stream = true;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        ctx->ncu = ncu;
    // fused-sort key counts: zero once here, cleared again by every sort
    const size_t fs_bytes = sizeof(int) * kFsMaxKeys;
    if (hipMalloc((void **)&ctx->fs, fs_bytes) != hipSuccess ||
        hipMemsetAsync(ctx->fs, 0, fs_bytes, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        if (ctx->fs) hipFree(ctx->fs);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return SE2P_ERR_HIP;
    }
    for (auto &stage : ctx->ev)
        for (auto &visit : stage)
            for (hipEvent_t &e : visit) hipEventCreate(&e);
    // this device's address of the no-allocation entry's fixed scratch; on
    // failure that entry reports it and every other entry works as before
    if (noalloc_scratch_init(ctx) != 0) ctx->scratch = nullptr;
    ctx->err.clear();
    *out = ctx;
    return SE2P_OK;
}

int se2p_destroy(se2p_ctx *ctx)
{
    if (!ctx) return SE2P_OK;
    if (ctx->device == SE2P_HOST_ONLY) {
        delete ctx;
        return SE2P_OK;
    }
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto *pool : {&ctx->slots, &ctx->gslots})   // a group left open may still run there
        for (auto &sl : *pool)
            if (sl.st) hipStreamSynchronize(sl.st);
    for (auto &kv : ctx->plans) {
        FftPlans &p = kv.second;
        if (p.fwd) rocfft_plan_destroy(p.fwd);
        if (p.inv) rocfft_plan_destroy(p.inv);
        if (p.fwd_info) rocfft_execution_info_destroy(p.fwd_info);
        if (p.inv_info) rocfft_execution_info_destroy(p.inv_info);
        if (p.work) hipFree(p.work);
    }
    free_kept_real(ctx);
    free_kept_k(ctx);
    if (ctx->ws) hipFree(ctx->ws);
    if (ctx->drv) hipFree(ctx->drv);
    if (ctx->io) hipFree(ctx->io);
    if (ctx->fs) hipFree(ctx->fs);
    for (auto *pool : {&ctx->slots, &ctx->gslots})
        for (auto &sl : *pool) {
            if (sl.ws) hipFree(sl.ws);
            if (sl.fs) hipFree(sl.fs);
            if (sl.done) hipEventDestroy(sl.done);
            if (sl.st) hipStreamDestroy(sl.st);
        }
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->switch_ev) hipEventDestroy(ctx->switch_ev);
    if (ctx->group_ev) hipEventDestroy(ctx->group_ev);
    for (auto &stage : ctx->ev)
        for (auto &visit : stage)
            for (hipEvent_t &e : visit)
                if (e) hipEventDestroy(e);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return SE2P_OK;
}

const char *se2p_last_error(const se2p_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int se2p_set_stream(se2p_ctx *ctx, void *s)
{
    if (!ctx) return SE2P_ERR_ARG;
    if (ctx->device == SE2P_HOST_ONLY) return fail(ctx, SE2P_ERR_HIP, "host-only context");
    if (s && (hipStream_t)s == ctx->stream) return SE2P_OK;
    if (!s && ctx->own_stream) return SE2P_OK;
    if (int rc = refuse_in_group(ctx, "binding another stream (se2p_set_stream)")) return rc;
    SE2P_HIP(hipSetDevice(ctx->device));
    // The context's workspace, sort state and kept buffers are shared by all
    // its calls: work enqueued on the new stream must not overlap what is
    // still pending on the old one, so the new stream waits for an event
    // recorded on the old one (no host synchronisation).  While either stream
    // is being captured into a graph no such link is made (it would join the
    // capture or wait on an event from outside it): the capturing caller
    // orders the streams.
    const hipStream_t old = ctx->stream;
    hipStream_t next = (hipStream_t)s;
    if (!s) SE2P_HIP(hipStreamCreate(&next));
    hipStreamCaptureStatus cs_old = hipStreamCaptureStatusNone, cs_new = hipStreamCaptureStatusNone;
    if (old) SE2P_HIP(hipStreamIsCapturing(old, &cs_old));
    SE2P_HIP(hipStreamIsCapturing(next, &cs_new));
    if (old && cs_old == hipStreamCaptureStatusNone && cs_new == hipStreamCaptureStatusNone) {
        if (!ctx->switch_ev)
            SE2P_HIP(hipEventCreateWithFlags(&ctx->switch_ev, hipEventDisableTiming));
        SE2P_HIP(hipEventRecord(ctx->switch_ev, old));
        SE2P_HIP(hipStreamWaitEvent(next, ctx->switch_ev, 0));
    }
    // a stream the context created is released once its pending work is done
    if (ctx->own_stream && old) SE2P_HIP(hipStreamDestroy(old));
    ctx->stream = next;
    ctx->own_stream = !s;
    return SE2P_OK;
}

void *se2p_get_stream(se2p_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int se2p_set_sort_mode(se2p_ctx *ctx, int mode)
{
    if (!ctx) return SE2P_ERR_ARG;
    if (mode & ~(SE2P_SORT_RADIX | SE2P_SORT_BOX_KERNEL))
        return fail(ctx, SE2P_ERR_ARG, "unknown sort mode bits 0x%x", mode);
    ctx->sort_mode = mode;
    return SE2P_OK;
}

int se2p_set_adjoint_passes(se2p_ctx *ctx, int passes)
{
    if (!ctx) return SE2P_ERR_ARG;
    if (passes != 1 && passes != 2) return fail(ctx, SE2P_ERR_ARG, "passes must be 1 or 2 (got %d)", passes);
    ctx->split_adjoint = passes == 2;
    return SE2P_OK;
}

int se2p_set_async(se2p_ctx *ctx, int a)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->async = a != 0;
    return SE2P_OK;
}

int se2p_set_host_pointers(se2p_ctx *ctx, int enable)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->host_ptrs = enable != 0;
    return SE2P_OK;
}

int se2p_set_keep(se2p_ctx *ctx, int enable)
{
    if (!ctx) return SE2P_ERR_ARG;
    if (enable == 0)
        if (int rc = refuse_in_group(ctx, "se2p_set_keep(ctx, 0)")) return rc;
    if (enable == 0 && (ctx->kept_real || ctx->kept_k))
        if (int rc = refuse_if_frozen(ctx, "freeing the kept forward state", 0)) return rc;
    ctx->keep = enable != 0;
    if (!ctx->keep && ctx->device != SE2P_HOST_ONLY) {
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
        free_kept_real(ctx);
        free_kept_k(ctx);
    }
    return SE2P_OK;
}

int se2p_set_spread_chunk(se2p_ctx *ctx, int chunk)
{
    if (!ctx) return SE2P_ERR_ARG;
    if (chunk != 0 && (chunk < SE2P_SPREAD_CHUNK_MIN || chunk > SE2P_SPREAD_CHUNK_MAX || chunk % 16 != 0))
        return fail(ctx, SE2P_ERR_ARG, "spread chunk %d: 0, or a multiple of 16 in [%d, %d]", chunk,
                    SE2P_SPREAD_CHUNK_MIN, SE2P_SPREAD_CHUNK_MAX);
    if (chunk != ctx->spread_chunk)
        if (int rc = refuse_in_group(ctx, "changing the spreading split")) return rc;
    if (chunk != ctx->spread_chunk && ctx->kept_k)
        if (int rc = refuse_if_frozen(ctx, "changing the split (it frees the kept state)", 0))
            return rc;
    if (chunk != ctx->spread_chunk && ctx->device != SE2P_HOST_ONLY) {
        // kept Fourier-space state carries the old work split
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
        free_kept_k(ctx);
    }
    ctx->spread_chunk = chunk;
    return SE2P_OK;
}

int se2p_set_frozen(se2p_ctx *ctx, int enable)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->frozen = enable != 0;
    return SE2P_OK;
}

int se2p_set_timing(se2p_ctx *ctx, int enable)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->timing = enable != 0;
    return SE2P_OK;
}

int se2p_stage_times(se2p_ctx *ctx, float *ms, int n)
{
    if (!ctx || !ms) return SE2P_ERR_ARG;
    for (int i = 0; i < n && i < ST_COUNT; ++i) ms[i] = ctx->timing ? ctx->stage_ms[i] : -1.0f;
    return SE2P_OK;
}

int se2p_group_begin(se2p_ctx *ctx)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->err.clear();
    if (ctx->group >= 0) return fail(ctx, SE2P_ERR_ARG, "a group is already open");
    if (int rc = set_device(ctx)) return rc;
    if (!ctx->group_ev) SE2P_HIP(hipEventCreateWithFlags(&ctx->group_ev, hipEventDisableTiming));
    SE2P_HIP(hipEventRecord(ctx->group_ev, ctx->stream));
    ctx->group = 0;
    return SE2P_OK;
}

int se2p_group_end(se2p_ctx *ctx)
{
    if (!ctx) return SE2P_ERR_ARG;
    ctx->err.clear();
    if (ctx->group < 0) return fail(ctx, SE2P_ERR_ARG, "no group is open");
    const int used = ctx->group;
    ctx->group = -1;
    SE2P_HIP(hipSetDevice(ctx->device));
    for (int k = 1; k < used; ++k)   // the first call ran on the context stream
        SE2P_HIP(hipStreamWaitEvent(ctx->stream, ctx->gslots[(size_t)k - 1].done, 0));
    if (!ctx->async) SE2P_HIP(hipStreamSynchronize(ctx->stream));
    return SE2P_OK;
}

int se2p_slp_real(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                  const double *f, double xi, int nside_x, int nside_y, double Lx, double Ly,
                  double *u)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && !f) return fail(ctx, SE2P_ERR_ARG, "null density pointer");
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(u, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    u = io.out(u);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_sum(ctx, 0, g, f, nullptr, 1, 2 * ns, 2 * nt, u, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_real(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                  const double *f, const double *n, double xi, int nside_x, int nside_y,
                  double Lx, double Ly, double *u)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    io.want(u, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    u = io.out(u);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_sum(ctx, 1, g, f, n, 1, 2 * ns, 2 * nt, u, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_real_noalloc(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, const double *n, double xi, int nside_x,
                          int nside_y, double Lx, double Ly, double *u)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = refuse_in_group(ctx, "se2p_dlp_real_noalloc")) return rc;
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (ctx->host_ptrs)
        return fail(ctx, SE2P_ERR_ARG,
                    "se2p_dlp_real_noalloc takes device pointers only (host-pointer mode "
                    "stages arrays in allocated memory)");
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    size_t need = 0;
    if (int rc = real_scratch_bytes(ctx, 1, g, &need)) return rc;
    if (need > (size_t)SE2P_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, SE2P_ERR_ARG,
                    "se2p_dlp_real_noalloc: Ns = %lld, Nt = %lld on %d x %d cells need %zu bytes "
                    "of scratch; the fixed scratch holds %zu (use se2p_dlp_real)",
                    (long long)ns, (long long)nt, nside_x, nside_y, need,
                    (size_t)SE2P_NOALLOC_SCRATCH_BYTES);
    if (int rc = set_device(ctx)) return rc;
    if (!ctx->scratch)
        return fail(ctx, SE2P_ERR_HIP, "no fixed scratch address on device %d", ctx->device);
    if (int rc = real_sum_fixed(ctx, 1, g, f, n, u, ctx->scratch, SE2P_NOALLOC_SCRATCH_BYTES))
        return rc;
    return finish(ctx);
}

int se2p_slp_kspace(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                    double xi, double eta, const double *f, int Mx, int My, double Lx, double Ly,
                    double w, int P, double *u)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && !f) return fail(ctx, SE2P_ERR_ARG, "null density pointer");
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(u, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    u = io.out(u);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_sum(ctx, 0, g, f, nullptr, 1, 2 * ns, 2 * nt, u)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_kspace(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                    double xi, double eta, const double *f, const double *n, int Mx, int My,
                    double Lx, double Ly, double w, int P, double *u)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    io.want(u, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    u = io.out(u);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_sum(ctx, 1, g, f, n, 1, 2 * ns, 2 * nt, u)) return rc;
    return io.done(finish(ctx));
}

static int check_batch(se2p_ctx *ctx, int nrhs, int64_t f_stride, int64_t ns, int64_t u_stride,
                       int64_t nt)
{
    if (nrhs < 1) return fail(ctx, SE2P_ERR_ARG, "nrhs must be >= 1");
    if (f_stride < 2 * ns || u_stride < 2 * nt)
        return fail(ctx, SE2P_ERR_ARG, "strides must be at least 2*Ns (densities) and 2*Nt (outputs)");
    return 0;
}

int se2p_slp_real_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, int nrhs, int64_t f_stride, double xi,
                          int nside_x, int nside_y, double Lx, double Ly, double *u,
                          int64_t u_stride)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && !f) return fail(ctx, SE2P_ERR_ARG, "null density pointer");
    if (int rc = check_batch(ctx, nrhs, f_stride, ns, u_stride, nt)) return rc;
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 8 * ((size_t)(nrhs - 1) * f_stride + 2 * ns));
    io.want(u, 8 * ((size_t)(nrhs - 1) * u_stride + 2 * nt));
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    u = io.out(u);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_sum(ctx, 0, g, f, nullptr, nrhs, f_stride, u_stride, u, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_real_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, const double *n, int nrhs, int64_t f_stride,
                          double xi, int nside_x, int nside_y, double Lx, double Ly, double *u,
                          int64_t u_stride)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    if (int rc = check_batch(ctx, nrhs, f_stride, ns, u_stride, nt)) return rc;
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 8 * ((size_t)(nrhs - 1) * f_stride + 2 * ns));
    io.want(n, 16 * (size_t)ns);
    io.want(u, 8 * ((size_t)(nrhs - 1) * u_stride + 2 * nt));
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    u = io.out(u);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_sum(ctx, 1, g, f, n, nrhs, f_stride, u_stride, u, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_slp_kspace_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                            int64_t nt, double xi, double eta, const double *f, int nrhs,
                            int64_t f_stride, int Mx, int My, double Lx, double Ly, double w, int P,
                            double *u, int64_t u_stride)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && !f) return fail(ctx, SE2P_ERR_ARG, "null density pointer");
    if (int rc = check_batch(ctx, nrhs, f_stride, ns, u_stride, nt)) return rc;
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 8 * ((size_t)(nrhs - 1) * f_stride + 2 * ns));
    io.want(u, 8 * ((size_t)(nrhs - 1) * u_stride + 2 * nt));
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    u = io.out(u);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_sum(ctx, 0, g, f, nullptr, nrhs, f_stride, u_stride, u)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_kspace_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                            int64_t nt, double xi, double eta, const double *f, const double *n,
                            int nrhs, int64_t f_stride, int Mx, int My, double Lx, double Ly,
                            double w, int P, double *u, int64_t u_stride)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    if (int rc = check_batch(ctx, nrhs, f_stride, ns, u_stride, nt)) return rc;
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 8 * ((size_t)(nrhs - 1) * f_stride + 2 * ns));
    io.want(n, 16 * (size_t)ns);
    io.want(u, 8 * ((size_t)(nrhs - 1) * u_stride + 2 * nt));
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    u = io.out(u);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_sum(ctx, 1, g, f, n, nrhs, f_stride, u_stride, u)) return rc;
    return io.done(finish(ctx));
}

static int check_box(se2p_ctx *ctx, const se2p_box_desc &b, int i, bool dlp, bool kspace)
{
    int rc = check_common(ctx, b.psrc, b.ns, b.ptar, b.nt, b.u);
    if (!rc && b.ns > 0 && (!b.f || (dlp && !b.n))) rc = fail(ctx, SE2P_ERR_ARG, "null density pointer");
    if (!rc && !kspace) rc = check_real(ctx, b.xi, b.nside_x, b.nside_y, b.Lx, b.Ly);
    if (!rc && kspace) {
        KGeom g{b.psrc, b.ptar, b.ns, b.nt, b.xi, b.eta, b.Lx, b.Ly, b.w, b.Mx, b.My, b.P};
        rc = check_kspace_params(ctx, g);
    }
    if (rc) ctx->err = "box " + std::to_string(i) + ": " + ctx->err;
    return rc;
}

// Restores the context's keep flag: multi-box calls keep no forward state
// (their boxes run on several streams).
struct KeepOff {
    se2p_ctx *ctx;
    bool was;
    explicit KeepOff(se2p_ctx *c) : ctx(c), was(c->keep) { c->keep = false; }
    ~KeepOff() { ctx->keep = was; }
};

static int kspace_multibox(se2p_ctx *ctx, int kind, const se2p_box_desc *boxes, int nbox)
{
    KeepOff no_keep(ctx);
    return kspace_sum_boxes(ctx, kind, boxes, nbox);
}

static int multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox, int kind, bool kspace)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = refuse_in_group(ctx, "a multi-box call")) return rc;
    if (nbox < 0 || (nbox > 0 && !boxes)) return fail(ctx, SE2P_ERR_ARG, "bad box list");
    for (int i = 0; i < nbox; ++i)
        if (int rc = check_box(ctx, boxes[i], i, kind == 1, kspace)) return rc;
    if (int rc = set_device(ctx)) return rc;
    HostIo io(ctx);
    std::vector<se2p_box_desc> staged;
    if (ctx->host_ptrs && nbox > 0) {
        for (int i = 0; i < nbox; ++i) {
            const se2p_box_desc &b = boxes[i];
            io.want(b.psrc, 16 * (size_t)b.ns);
            io.want(b.ptar, 16 * (size_t)b.nt);
            io.want(b.f, 16 * (size_t)b.ns);
            if (kind == 1) io.want(b.n, 16 * (size_t)b.ns);
            io.want(b.u, 16 * (size_t)b.nt);
        }
        if (int rc = io.map()) return rc;
        staged.assign(boxes, boxes + nbox);
        for (se2p_box_desc &b : staged) {
            b.psrc = io.in(b.psrc);
            b.ptar = io.in(b.ptar);
            b.f = io.in(b.f);
            if (kind == 1) b.n = io.in(b.n);
            b.u = io.out(b.u);
        }
        boxes = staged.data();
    }
    if (!kspace) {
        if (nbox > 0)
            if (int rc = real_sum_multibox(ctx, kind, boxes, nbox)) return rc;
    } else if (nbox > 0) {
        if (int rc = kspace_multibox(ctx, kind, boxes, nbox)) return rc;
    }
    return io.done(finish(ctx));
}

int se2p_slp_real_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox)
{
    return multibox(ctx, boxes, nbox, 0, false);
}

int se2p_slp_kspace_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox)
{
    return multibox(ctx, boxes, nbox, 0, true);
}

int se2p_dlp_real_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox)
{
    return multibox(ctx, boxes, nbox, 1, false);
}

int se2p_dlp_kspace_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox)
{
    return multibox(ctx, boxes, nbox, 1, true);
}

int se2p_slp_real_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                           int64_t nt, const double *f, double xi, int nside_x, int nside_y,
                           double Lx, double Ly, const double *du, double *dpsrc, double *dptar,
                           double *df)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_adjoint(ctx, psrc, ns, ptar, nt, f, nullptr, false, du)) return rc;
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(du, 16 * (size_t)nt);
    io.want(dpsrc, 16 * (size_t)ns);
    io.want(dptar, 16 * (size_t)nt);
    io.want(df, 16 * (size_t)ns);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    du = io.in(du);
    dpsrc = io.out(dpsrc);
    dptar = io.out(dptar);
    df = io.out(df);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_adjoint(ctx, 0, g, f, nullptr, du, dpsrc, dptar, df, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_real_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                           int64_t nt, const double *f, const double *n, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, const double *du, double *dpsrc,
                           double *dptar, double *df, double *dn)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_adjoint(ctx, psrc, ns, ptar, nt, f, n, true, du)) return rc;
    if (int rc = check_real(ctx, xi, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    io.want(du, 16 * (size_t)nt);
    io.want(dpsrc, 16 * (size_t)ns);
    io.want(dptar, 16 * (size_t)nt);
    io.want(df, 16 * (size_t)ns);
    io.want(dn, 16 * (size_t)ns);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    du = io.in(du);
    dpsrc = io.out(dpsrc);
    dptar = io.out(dptar);
    df = io.out(df);
    dn = io.out(dn);
    RealGeom g{psrc, ptar, ns, nt, xi, Lx, Ly, nside_x, nside_y};
    if (int rc = real_adjoint(ctx, 1, g, f, n, du, dpsrc, dptar, df, dn)) return rc;
    return io.done(finish(ctx));
}

int se2p_slp_kspace_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                             int64_t nt, double xi, double eta, const double *f, int Mx, int My,
                             double Lx, double Ly, double w, int P, const double *du,
                             double *dpsrc, double *dptar, double *df)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_adjoint(ctx, psrc, ns, ptar, nt, f, nullptr, false, du)) return rc;
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(du, 16 * (size_t)nt);
    io.want(dpsrc, 16 * (size_t)ns);
    io.want(dptar, 16 * (size_t)nt);
    io.want(df, 16 * (size_t)ns);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    du = io.in(du);
    dpsrc = io.out(dpsrc);
    dptar = io.out(dptar);
    df = io.out(df);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_adjoint(ctx, 0, g, f, nullptr, du, dpsrc, dptar, df, nullptr)) return rc;
    return io.done(finish(ctx));
}

int se2p_dlp_kspace_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                             int64_t nt, double xi, double eta, const double *f, const double *n,
                             int Mx, int My, double Lx, double Ly, double w, int P,
                             const double *du, double *dpsrc, double *dptar, double *df,
                             double *dn)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_adjoint(ctx, psrc, ns, ptar, nt, f, n, true, du)) return rc;
    KGeom g{psrc, ptar, ns, nt, xi, eta, Lx, Ly, w, Mx, My, P};
    if (int rc = check_kspace_params(ctx, g)) return rc;
    if (int rc = set_device(ctx)) return rc;
    GroupCall gc(ctx);
    if (int rc = gc.enter()) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    io.want(du, 16 * (size_t)nt);
    io.want(dpsrc, 16 * (size_t)ns);
    io.want(dptar, 16 * (size_t)nt);
    io.want(df, 16 * (size_t)ns);
    io.want(dn, 16 * (size_t)ns);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    du = io.in(du);
    dpsrc = io.out(dpsrc);
    dptar = io.out(dptar);
    df = io.out(df);
    dn = io.out(dn);
    g.psrc = psrc;
    g.ptar = ptar;
    if (int rc = kspace_adjoint(ctx, 1, g, f, n, du, dpsrc, dptar, df, dn)) return rc;
    return io.done(finish(ctx));
}

int se2p_select_params(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                       se2p_params *p)
{
    if (!p || (kind != 0 && kind != 1) || ns < 0 || nt < 0 || ns + nt < 2 || !(Lx > 0) || !(Ly > 0) ||
        !(Q >= 1) || !std::isfinite(Q))
        return SE2P_ERR_ARG;
    if (p->P < 0 || (p->P % 2) != 0) return SE2P_ERR_P;
    if (p->fixed_grid && (p->M < 2 || p->M % 2 != 0)) return SE2P_ERR_GRID;
    return select_params(kind, ns, nt, Q, Lx, Ly, p);
}

static int ensure_driver_buffer(se2p_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->drv_bytes) return 0;
    if (int rc = refuse_if_frozen(ctx, "growing the driver buffers", bytes)) return rc;
    if (ctx->drv) {
        SE2P_HIP(hipStreamSynchronize(ctx->stream));
        SE2P_HIP(hipFree(ctx->drv));
        ctx->drv = nullptr;
        ctx->drv_bytes = 0;
    }
    SE2P_HIP(hipMalloc((void **)&ctx->drv, bytes));
    ctx->drv_bytes = bytes;
    return 0;
}

// Driver-level device buffers: the wrapped points, the real- and
// Fourier-space outputs and the reduction scratch of launch_driver_sums.
struct DrvBufs {
    double *src, *tar, *ur, *uk, *partial, *sums;
};

static int driver_buffers(se2p_ctx *ctx, int64_t ns, int64_t nt, bool same, DrvBufs *b)
{
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t s16 = up(16 * (size_t)(ns > 0 ? ns : 1)), t16 = up(16 * (size_t)(nt > 0 ? nt : 1));
    const int64_t nb = (ns + 255) / 256;
    const size_t need = s16 + (same ? 0 : t16) + 2 * t16 + up(24 * (size_t)(nb + 1)) + up(24);
    if (int rc = ensure_driver_buffer(ctx, need)) return rc;
    size_t o = 0;
    b->src = (double *)(ctx->drv + o); o += s16;
    b->tar = b->src;
    if (!same) { b->tar = (double *)(ctx->drv + o); o += t16; }
    b->ur = (double *)(ctx->drv + o); o += t16;
    b->uk = (double *)(ctx->drv + o); o += t16;
    b->partial = (double *)(ctx->drv + o); o += up(24 * (size_t)(nb + 1));
    b->sums = (double *)(ctx->drv + o);
    return 0;
}

// Shared checks of the driver-level entries.
static int check_driver(se2p_ctx *ctx, int64_t ns, int64_t nt, double Lx, double Ly,
                        const se2p_params *prm)
{
    if (!prm) return fail(ctx, SE2P_ERR_ARG, "null parameter struct");
    if (ns + nt < 2) return fail(ctx, SE2P_ERR_ARG, "need at least two points in total");
    if (!(Lx > 0) || !(Ly > 0) || !std::isfinite(Lx) || !std::isfinite(Ly))
        return fail(ctx, SE2P_ERR_PARAM, "Lx and Ly must be positive and finite");
    if (prm->P < 0 || (prm->P % 2) != 0) return fail(ctx, SE2P_ERR_P, "P must be even");
    if (prm->fixed_grid && (prm->M < 2 || prm->M % 2 != 0))
        return fail(ctx, SE2P_ERR_GRID, "fixed-grid mode needs an even grid M >= 2 (got %d)", prm->M);
    return 0;
}

// Shared body of the two driver entry points.
static int ewald_full(se2p_ctx *ctx, int kind, const double *psrc, int64_t ns, const double *ptar,
                      int64_t nt, const double *f, const double *n, double Lx, double Ly,
                      se2p_params *prm, double *u, double *ur_out, double *uk_out)
{
    if (int rc = refuse_in_group(ctx, "a driver call")) return rc;
    if (int rc = check_driver(ctx, ns, nt, Lx, Ly, prm)) return rc;
    if (int rc = set_device(ctx)) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    io.want(u, 16 * (size_t)nt);
    io.want(ur_out, 16 * (size_t)nt);
    io.want(uk_out, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    u = io.out(u);
    ur_out = io.out(ur_out);
    uk_out = io.out(uk_out);
    hipStream_t st = ctx->stream;
    const bool same = (psrc == ptar) && (ns == nt);
    DrvBufs b;
    if (int rc = driver_buffers(ctx, ns, nt, same, &b)) return rc;
    double *ur = ur_out ? ur_out : b.ur;
    double *uk = uk_out ? uk_out : b.uk;

    // Q = sum(f.^2) + 1 and, for the DLP, sum((f.n) x) over the caller's points
    launch_driver_sums(st, f, kind == 1 ? n : nullptr, psrc, ns, b.partial, b.sums);
    double hs[3];
    SE2P_HIP(hipMemcpyAsync(hs, b.sums, sizeof(hs), hipMemcpyDeviceToHost, st));
    SE2P_HIP(hipStreamSynchronize(st));
    if (select_params(kind, ns, nt, hs[0] + 1.0, Lx, Ly, prm) != 0)
        return fail(ctx, SE2P_ERR_PARAM, "no parameters meet tol %g on the fixed %d grid", prm->tol,
                    prm->M);

    launch_wrap(st, psrc, ns, Lx, Ly, b.src);
    if (!same) launch_wrap(st, ptar, nt, Lx, Ly, b.tar);
    RealGeom rg{b.src, b.tar, ns, nt, prm->xi, Lx, Ly, prm->nside_x, prm->nside_y};
    KGeom kg{b.src, b.tar, ns, nt, prm->xi, prm->eta, Lx, Ly, prm->w, prm->Mx, prm->My, prm->P};
    if (int rc = check_kspace_params(ctx, kg)) return rc;
    // the two sums read the wrapped points and write uk and ur only
    if (int rc = sums_side_by_side(
            ctx, [&] { return kspace_sum(ctx, kind, kg, f, n, 1, 2 * ns, 2 * nt, uk); },
            [&] { return real_sum(ctx, kind, rg, f, n, 1, 2 * ns, 2 * nt, ur, nullptr); }))
        return rc;
    launch_compose(st, ur, uk, nt, kind, b.sums, Lx * Ly, u);
    return io.done(finish(ctx));
}

// ---- autotuner -----------------------------------------------------------------
namespace {
struct TuneKey {
    int dev, kind, P;
    int64_t ns, nt;
    int same;
    double Lx, Ly, tol;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(dev, kind, P, ns, nt, same, Lx, Ly, tol) <
               std::tie(o.dev, o.kind, o.P, o.ns, o.nt, o.same, o.Lx, o.Ly, o.tol);
    }
};
std::mutex g_tune_mu;
struct TuneChoice {
    double Nb;   // points per cell
    int chunk;   // spreading chunk
    int P, M;    // support and grid override (M = 0: the drivers' grid)
};
std::map<TuneKey, TuneChoice> g_tune_nb;   // key -> choice

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Restores the context's timing flag when tuning ends.
struct TimingOff {
    se2p_ctx *ctx;
    bool was;
    explicit TimingOff(se2p_ctx *c) : ctx(c), was(c->timing) { c->timing = false; }
    ~TimingOff() { ctx->timing = was; }
};

// One geometry to tune: wrapped device points, densities, Q = sum f^2 + 1,
// and output scratch (2 x nt doubles each).
struct TuneGeom {
    int kind;
    const double *src, *tar;
    int64_t ns, nt;
    const double *f, *n;
    double Lx, Ly, Q, tol;
    double *ur, *uk;
};

enum TuneStage : unsigned { TUNE_NB = 1, TUNE_CHUNK = 2, TUNE_SUPPORT = 4 };
constexpr int kTuneReps = 3, kTuneBlocks = 3;
constexpr double kTuneMaxGrid = 4096.0 * 4096.0;
}  // namespace

// Driver parameters of geometry tg for points per cell Nb, support P and grid
// override M (0: the drivers' grid).
static void tune_params(const TuneGeom &tg, double Nb, int P, int M, se2p_params *c)
{
    *c = se2p_params{};
    c->P = P;
    c->Nb = Nb;
    c->tol = tg.tol;
    c->M = M;
    select_params(tg.kind, tg.ns, tg.nt, tg.Q, tg.Lx, tg.Ly, c);
    c->M = M;
}

// Best of kTuneBlocks blocks of kTuneReps back-to-back real + Fourier steps
// at parameters c, after one warm-up step (it sizes the workspace and makes
// the FFT plans); HIP events on the context stream.
static int tune_time(se2p_ctx *ctx, const TuneGeom &tg, const se2p_params &c, EventPair &ev, double *t)
{
    hipStream_t st = ctx->stream;
    const RealGeom rg{tg.src, tg.tar, tg.ns, tg.nt, c.xi, tg.Lx, tg.Ly, c.nside_x, c.nside_y};
    const KGeom kg{tg.src, tg.tar, tg.ns, tg.nt, c.xi, c.eta, tg.Lx, tg.Ly, c.w, c.Mx, c.My, c.P};
    if (int rc = check_kspace_params(ctx, kg)) return rc;
    // the step the driver entries run: both sums side by side
    auto step = [&]() -> int {
        return sums_side_by_side(
            ctx, [&] { return kspace_sum(ctx, tg.kind, kg, tg.f, tg.n, 1, 2 * tg.ns, 2 * tg.nt, tg.uk); },
            [&] { return real_sum(ctx, tg.kind, rg, tg.f, tg.n, 1, 2 * tg.ns, 2 * tg.nt, tg.ur, nullptr); });
    };
    if (int rc = step()) return rc;
    SE2P_HIP(hipStreamSynchronize(st));
    *t = std::numeric_limits<double>::infinity();
    for (int blk = 0; blk < kTuneBlocks; ++blk) {
        SE2P_HIP(hipEventRecord(ev.a, st));
        for (int r = 0; r < kTuneReps; ++r)
            if (int rc = step()) return rc;
        SE2P_HIP(hipEventRecord(ev.b, st));
        SE2P_HIP(hipEventSynchronize(ev.b));
        float ms = 0.0f;
        SE2P_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        *t = std::min(*t, (double)ms / kTuneReps);
    }
    return 0;
}

// The tuning stages of one geometry (see se2p_autotune).  Pin: the caller's
// support (0: the drivers' 24, and stage 3 may change it).  Sets the chosen
// spreading chunk on the context when stage 2 runs.
static int tune_geometry(se2p_ctx *ctx, const TuneGeom &tg, int Pin, unsigned stages,
                         se2p_params *out, int *chunk_out, se2p_tune_info *info)
{
    const int P0 = Pin > 0 ? Pin : 24;   // the drivers' default support
    if (info) info->P = P0;
    TimingOff quiet(ctx);
    KeepOff no_keep(ctx);
    EventPair ev;
    SE2P_HIP(hipEventCreate(&ev.a));
    SE2P_HIP(hipEventCreate(&ev.b));

    // 1. the cell size: points per cell around the drivers' default, one
    //    candidate per cell grid
    const double npts = (double)(tg.ns + tg.nt);
    const double Nb0 = std::min(P0 * std::round(std::log2(npts)), npts / 4);
    std::vector<se2p_params> cand;
    if (stages & TUNE_NB) {
        for (int k = -6; k <= 2 && (int)cand.size() < SE2P_TUNE_MAX; ++k) {
            se2p_params c;
            tune_params(tg, Nb0 * std::pow(2.0, 0.5 * k), P0, 0, &c);
            if ((double)c.Mx * (double)c.My > kTuneMaxGrid) continue;
            bool dup = false;
            for (const se2p_params &o : cand) dup = dup || (o.nside_x == c.nside_x && o.nside_y == c.nside_y);
            if (!dup) cand.push_back(c);
        }
        if (cand.empty()) return fail(ctx, SE2P_ERR_PARAM, "no candidate grid within 4096^2 points");
    } else {
        se2p_params c;
        tune_params(tg, 0.0, P0, 0, &c);   // the drivers' default cells
        cand.push_back(c);
    }
    if (stages & TUNE_CHUNK)
        if (int rc = se2p_set_spread_chunk(ctx, 0)) return rc;
    int best = 0;
    double best_ms = std::numeric_limits<double>::infinity();
    for (size_t i = 0; i < cand.size(); ++i) {
        const se2p_params &c = cand[i];
        double t;
        if (int rc = tune_time(ctx, tg, c, ev, &t)) return rc;
        if (info && i < SE2P_TUNE_MAX) {
            info->Nb[i] = c.Nb;
            info->nside_x[i] = c.nside_x;
            info->Mx[i] = c.Mx;
            info->ms[i] = t;
        }
        if (t < best_ms) {
            best_ms = t;
            best = (int)i;
        }
    }
    if (info) {
        info->ncand = (int)cand.size();
        info->best = best;
    }
    // 2. the spreading work split for that cell size (0: the library's
    //    default, 256 candidates for one density; timed above)
    int best_chunk = 0;
    if (stages & TUNE_CHUNK) {
        const int chunks[SE2P_TUNE_CHUNKS] = {0, 128, 512};
        double chunk_best = best_ms;
        for (int k = 0; k < SE2P_TUNE_CHUNKS; ++k) {
            double t = best_ms;
            if (k > 0) {
                if (int rc = se2p_set_spread_chunk(ctx, chunks[k])) return rc;
                if (int rc = tune_time(ctx, tg, cand[(size_t)best], ev, &t)) return rc;
            }
            if (info) {
                info->chunk[k] = chunks[k];
                info->chunk_ms[k] = t;
            }
            if (t < chunk_best) {
                chunk_best = t;
                best_chunk = chunks[k];
            }
        }
        if (int rc = se2p_set_spread_chunk(ctx, best_chunk)) return rc;
        best_ms = chunk_best;
        if (info) info->spread_chunk = best_chunk;
    }
    se2p_params chosen = cand[(size_t)best];
    // 3. a narrower support on an oversampled grid: supports 16 and 20 on
    //    grids 1, 1.2 and 1.4 times the chosen one.  A candidate faster than
    //    the best so far is kept only if its Fourier-space sum differs from
    //    the chosen P = 24 one by at most tol/4 of the largest |ur + uk|: the
    //    Fourier-space sum converges as the support and the grid grow, so
    //    the difference bounds what the candidate adds to the error.
    if ((stages & TUNE_SUPPORT) && Pin == 0) {
        hipStream_t st = ctx->stream;
        const size_t nv = 2 * (size_t)tg.nt;
        std::vector<double> ur_h(nv), uk_ref(nv), uk_c(nv);
        {
            double t;
            if (int rc = tune_time(ctx, tg, chosen, ev, &t)) return rc;
        }
        SE2P_HIP(hipMemcpyAsync(ur_h.data(), tg.ur, nv * sizeof(double), hipMemcpyDeviceToHost, st));
        SE2P_HIP(hipMemcpyAsync(uk_ref.data(), tg.uk, nv * sizeof(double), hipMemcpyDeviceToHost, st));
        SE2P_HIP(hipStreamSynchronize(st));
        double scale = 0.0;
        for (size_t i = 0; i < nv; ++i) scale = std::max(scale, std::fabs(ur_h[i] + uk_ref[i]));
        const int Mb0 = std::min(2 * chosen.kinf, 10000);   // the drivers' grid per unit of A
        const int supports[2] = {16, 20};
        const double over[3] = {1.0, 1.2, 1.4};
        int np = 0;
        for (int ps : supports)
            for (double s : over) {
                const int M = 2 * (int)std::ceil(s * Mb0 / 2.0);
                se2p_params c;
                tune_params(tg, chosen.Nb, ps, M, &c);
                if ((double)c.Mx * (double)c.My > kTuneMaxGrid || c.Mx < ps / 2 + 2 || c.My < ps / 2 + 2)
                    continue;
                double t;
                if (int rc = tune_time(ctx, tg, c, ev, &t)) return rc;
                double dev = -1.0;
                if (t < best_ms) {
                    SE2P_HIP(hipMemcpyAsync(uk_c.data(), tg.uk, nv * sizeof(double), hipMemcpyDeviceToHost, st));
                    SE2P_HIP(hipStreamSynchronize(st));
                    double dmax = 0.0;
                    for (size_t i = 0; i < nv; ++i) dmax = std::max(dmax, std::fabs(uk_c[i] - uk_ref[i]));
                    dev = scale > 0 ? dmax / scale : dmax;
                    if (dev <= 0.25 * tg.tol) {
                        chosen = c;
                        best_ms = t;
                    }
                }
                if (info && np < SE2P_TUNE_PCAND) {
                    info->pc_P[np] = c.P;
                    info->pc_Mx[np] = c.Mx;
                    info->pc_ms[np] = t;
                    info->pc_dev[np] = dev;
                }
                ++np;
            }
        if (info) info->npc = std::min(np, SE2P_TUNE_PCAND);
    }
    SE2P_HIP(hipGetLastError());
    if (info) info->P = chosen.P;
    *out = chosen;
    if (chunk_out) *chunk_out = best_chunk;
    return 0;
}

static int autotune(se2p_ctx *ctx, int kind, const double *psrc, int64_t ns, const double *ptar,
                    int64_t nt, const double *f, const double *n, double Lx, double Ly,
                    se2p_params *prm, se2p_tune_info *info)
{
    if (int rc = refuse_in_group(ctx, "autotuning")) return rc;
    if (int rc = check_driver(ctx, ns, nt, Lx, Ly, prm)) return rc;
    if (!(prm->tol > 0) || !(prm->tol < 1))
        return fail(ctx, SE2P_ERR_PARAM, "tol must be in (0, 1)");
    if (prm->M != 0) return fail(ctx, SE2P_ERR_ARG, "M must be 0: the grid follows from the tuned cell size");
    if (prm->fixed_grid) return fail(ctx, SE2P_ERR_ARG, "the fixed-grid mode is not tuned");
    if (info) std::memset(info, 0, sizeof(*info));
    if (int rc = set_device(ctx)) return rc;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    io.want(f, 16 * (size_t)ns);
    io.want(n, 16 * (size_t)ns);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    f = io.in(f);
    n = io.in(n);
    hipStream_t st = ctx->stream;
    const bool same = (psrc == ptar) && (ns == nt);
    DrvBufs b;
    if (int rc = driver_buffers(ctx, ns, nt, same, &b)) return rc;
    launch_driver_sums(st, f, kind == 1 ? n : nullptr, psrc, ns, b.partial, b.sums);
    double hs[3];
    SE2P_HIP(hipMemcpyAsync(hs, b.sums, sizeof(hs), hipMemcpyDeviceToHost, st));
    SE2P_HIP(hipStreamSynchronize(st));
    const TuneGeom tg{kind, b.src, b.tar, ns, nt, f, n, Lx, Ly, hs[0] + 1.0, prm->tol, b.ur, b.uk};
    const TuneKey key{ctx->device, kind, prm->P, ns, nt, same ? 1 : 0, Lx, Ly, prm->tol};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_nb.find(key);
        if (it != g_tune_nb.end()) {
            tune_params(tg, it->second.Nb, it->second.P, it->second.M, prm);
            if (int rc = se2p_set_spread_chunk(ctx, it->second.chunk)) return rc;
            if (info) {
                info->cached = 1;
                info->P = prm->P;
                info->spread_chunk = it->second.chunk;
            }
            return io.done(SE2P_OK);
        }
    }
    launch_wrap(st, psrc, ns, Lx, Ly, b.src);
    if (!same) launch_wrap(st, ptar, nt, Lx, Ly, b.tar);
    int chunk = 0;
    if (int rc = tune_geometry(ctx, tg, prm->P, TUNE_NB | TUNE_CHUNK | TUNE_SUPPORT, prm, &chunk, info))
        return rc;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_nb[key] = TuneChoice{prm->Nb, chunk, prm->P, prm->M};
    return io.done(SE2P_OK);
}

// Smallest even length >= M whose prime factors are 2, 3, 5 and 7.
static int smooth_even_at_least(int M)
{
    for (int m = M + (M & 1);; m += 2) {
        int r = m;
        for (int f : {2, 3, 5, 7})
            while (r % f == 0) r /= f;
        if (r == 1) return m;
    }
}

// Multi-box tuning: each box's cell size (stage 1 of the single-box tuner at
// its drivers' P = 24), written into its descriptor, then the spreading work
// split of the packed Fourier-space call over all boxes.
static int autotune_boxes(se2p_ctx *ctx, int kind, se2p_box_desc *boxes, int nbox, double tol,
                          se2p_tune_boxes_info *info)
{
    if (int rc = refuse_in_group(ctx, "autotuning")) return rc;
    if (!(tol > 0) || !(tol < 1)) return fail(ctx, SE2P_ERR_PARAM, "tol must be in (0, 1)");
    if (nbox < 0 || (nbox > 0 && !boxes)) return fail(ctx, SE2P_ERR_ARG, "bad box list");
    if (ctx->host_ptrs) return fail(ctx, SE2P_ERR_ARG, "multi-box tuning needs device arrays");
    if (info) std::memset(info, 0, sizeof(*info));
    for (int i = 0; i < nbox; ++i) {
        const se2p_box_desc &d = boxes[i];
        if (int rc = check_common(ctx, d.psrc, d.ns, d.ptar, d.nt, d.u)) {
            ctx->err = "box " + std::to_string(i) + ": " + ctx->err;
            return rc;
        }
        if (d.ns > 0 && (!d.f || (kind == 1 && !d.n)))
            return fail(ctx, SE2P_ERR_ARG, "box %d: null density pointer", i);
        if (d.ns + d.nt < 2) return fail(ctx, SE2P_ERR_ARG, "box %d: need at least two points in total", i);
        if (!(d.Lx > 0) || !(d.Ly > 0) || !std::isfinite(d.Lx) || !std::isfinite(d.Ly))
            return fail(ctx, SE2P_ERR_PARAM, "box %d: Lx and Ly must be positive and finite", i);
    }
    if (int rc = set_device(ctx)) return rc;
    hipStream_t st = ctx->stream;
    EventPair ev;
    SE2P_HIP(hipEventCreate(&ev.a));
    SE2P_HIP(hipEventCreate(&ev.b));
    auto step = [&]() -> int {
        if (int rc = real_sum_multibox(ctx, kind, boxes, nbox)) return rc;
        return kspace_multibox(ctx, kind, boxes, nbox);
    };
    auto time_boxes = [&](double *t) -> int {
        if (int rc = step()) return rc;
        SE2P_HIP(hipStreamSynchronize(st));
        *t = std::numeric_limits<double>::infinity();
        for (int blk = 0; blk < kTuneBlocks; ++blk) {
            SE2P_HIP(hipEventRecord(ev.a, st));
            for (int r = 0; r < kTuneReps; ++r)
                if (int rc = step()) return rc;
            SE2P_HIP(hipEventRecord(ev.b, st));
            SE2P_HIP(hipEventSynchronize(ev.b));
            float ms = 0.0f;
            SE2P_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
            *t = std::min(*t, (double)ms / kTuneReps);
        }
        return 0;
    };
    TimingOff quiet(ctx);
    if (int rc = se2p_set_spread_chunk(ctx, 0)) return rc;
    // the drivers' parameters first, for the record
    std::vector<se2p_params> drv((size_t)nbox), tuned((size_t)nbox);
    std::vector<double> Q((size_t)nbox, 1.0);
    for (int i = 0; i < nbox; ++i) {
        se2p_box_desc &d = boxes[i];
        const bool same = d.psrc == d.ptar && d.ns == d.nt;
        DrvBufs b;
        if (int rc = driver_buffers(ctx, d.ns, d.nt, same, &b)) return rc;
        launch_driver_sums(st, d.f, kind == 1 ? d.n : nullptr, d.psrc, d.ns, b.partial, b.sums);
        double hs[3];
        SE2P_HIP(hipMemcpyAsync(hs, b.sums, sizeof(hs), hipMemcpyDeviceToHost, st));
        SE2P_HIP(hipStreamSynchronize(st));
        Q[(size_t)i] = hs[0] + 1.0;
        const TuneGeom tg{kind, d.psrc, d.ptar, d.ns, d.nt, d.f, d.n, d.Lx, d.Ly, Q[(size_t)i], tol, b.ur, b.uk};
        tune_params(tg, 0.0, 24, 0, &drv[(size_t)i]);
    }
    auto apply = [&](const std::vector<se2p_params> &ps) {
        for (int i = 0; i < nbox; ++i) {
            const se2p_params &c = ps[(size_t)i];
            se2p_box_desc &d = boxes[i];
            d.xi = c.xi;
            d.nside_x = c.nside_x;
            d.nside_y = c.nside_y;
            d.eta = c.eta;
            d.w = c.w;
            d.Mx = c.Mx;
            d.My = c.My;
            d.P = c.P;
        }
    };
    apply(drv);
    for (int i = 0; i < nbox; ++i)
        for (bool ks : {false, true})
            if (int rc = check_box(ctx, boxes[i], i, kind == 1, ks)) return rc;
    double t_drv = 0.0;
    if (int rc = time_boxes(&t_drv)) return rc;
    // each box's cell size, timed on its own
    for (int i = 0; i < nbox; ++i) {
        se2p_box_desc &d = boxes[i];
        tuned[(size_t)i] = drv[(size_t)i];
        if (d.ns == 0 || d.nt == 0) continue;
        const bool same = d.psrc == d.ptar && d.ns == d.nt;
        DrvBufs b;
        if (int rc = driver_buffers(ctx, d.ns, d.nt, same, &b)) return rc;
        const TuneGeom tg{kind, d.psrc, d.ptar, d.ns, d.nt, d.f, d.n, d.Lx, d.Ly, Q[(size_t)i], tol, b.ur, b.uk};
        // a cell size se2p_autotune chose for this geometry (P = 24) in this process
        bool cached = false;
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            const TuneKey key{ctx->device, kind, 24, d.ns, d.nt, same ? 1 : 0, d.Lx, d.Ly, tol};
            auto it = g_tune_nb.find(key);
            if (it != g_tune_nb.end()) {
                tune_params(tg, it->second.Nb, 24, 0, &tuned[(size_t)i]);
                cached = true;
            }
        }
        if (!cached)
            if (int rc = tune_geometry(ctx, tg, 24, TUNE_NB, &tuned[(size_t)i], nullptr, nullptr)) {
                ctx->err = "box " + std::to_string(i) + ": " + ctx->err;
                return rc;
            }
    }
    apply(tuned);
    for (int i = 0; i < nbox; ++i)
        for (bool ks : {false, true})
            if (int rc = check_box(ctx, boxes[i], i, kind == 1, ks)) return rc;
    // the spreading work split of the packed call
    const int chunks[SE2P_TUNE_CHUNKS] = {0, 128, 512};
    int best_chunk = 0;
    double best = std::numeric_limits<double>::infinity();
    for (int k = 0; k < SE2P_TUNE_CHUNKS; ++k) {
        if (int rc = se2p_set_spread_chunk(ctx, chunks[k])) return rc;
        double t;
        if (int rc = time_boxes(&t)) return rc;
        if (info) {
            info->chunk[k] = chunks[k];
            info->chunk_ms[k] = t;
        }
        if (t < best) {
            best = t;
            best_chunk = chunks[k];
        }
    }
    // keep the drivers' cells where the tuned ones are not faster as a whole
    std::vector<se2p_params> cur = tuned;
    const bool use_drv = !(best < t_drv);
    if (use_drv) {
        cur = drv;
        best_chunk = 0;
        best = t_drv;
    }
    apply(cur);
    if (int rc = se2p_set_spread_chunk(ctx, best_chunk)) return rc;
    // grids rounded up to lengths 2^a 3^b 5^c 7^d (rocFFT transforms them
    // without Bluestein stages): the same xi, cells and support on a finer
    // grid, so the Fourier-space truncation error can only fall
    std::vector<se2p_params> smooth = cur;
    int nround = 0;
    for (int i = 0; i < nbox; ++i) {
        const se2p_params &c = cur[(size_t)i];
        const int Mb = c.M > 0 ? c.M : std::min(2 * c.kinf, 10000);
        const int Ms = smooth_even_at_least(Mb);
        if (Ms == Mb) continue;
        const se2p_box_desc &d = boxes[i];
        const TuneGeom tg{kind, d.psrc, d.ptar, d.ns, d.nt, d.f, d.n, d.Lx, d.Ly, Q[(size_t)i], tol,
                          nullptr, nullptr};
        tune_params(tg, c.Nb, c.P, Ms, &smooth[(size_t)i]);
        if ((double)smooth[(size_t)i].Mx * smooth[(size_t)i].My > kTuneMaxGrid) smooth[(size_t)i] = c;
        else ++nround;
    }
    double t_smooth = -1.0;
    bool use_smooth = false;
    if (nround > 0) {
        apply(smooth);
        for (int i = 0; i < nbox; ++i)
            if (int rc = check_box(ctx, boxes[i], i, kind == 1, true)) return rc;
        if (int rc = time_boxes(&t_smooth)) return rc;
        use_smooth = t_smooth < best;
        if (use_smooth) best = t_smooth;
        else apply(cur);
    }
    if (info) {
        info->drivers_ms = t_drv;
        info->tuned_ms = best;
        info->spread_chunk = best_chunk;
        info->kept_drivers = use_drv && !use_smooth ? 1 : 0;
        info->smooth_grids = use_smooth ? nround : 0;
        info->smooth_ms = t_smooth;
    }
    SE2P_HIP(hipGetLastError());
    return finish(ctx);
}

int se2p_slp_ewald(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                   const double *f, double Lx, double Ly, se2p_params *prm, double *u, double *ur,
                   double *uk)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && !f) return fail(ctx, SE2P_ERR_ARG, "null density pointer");
    return ewald_full(ctx, 0, psrc, ns, ptar, nt, f, nullptr, Lx, Ly, prm, u, ur, uk);
}

int se2p_dlp_ewald(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                   const double *f, const double *n, double Lx, double Ly, se2p_params *prm,
                   double *u, double *ur, double *uk)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, u)) return rc;
    if (ns > 0 && (!f || !n)) return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    return ewald_full(ctx, 1, psrc, ns, ptar, nt, f, n, Lx, Ly, prm, u, ur, uk);
}

int se2p_autotune(se2p_ctx *ctx, int kind, const double *psrc, int64_t ns, const double *ptar,
                  int64_t nt, const double *f, const double *n, double Lx, double Ly,
                  se2p_params *prm, se2p_tune_info *info)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (kind != 0 && kind != 1) return fail(ctx, SE2P_ERR_ARG, "kind must be 0 (SLP) or 1 (DLP)");
    // the tuner writes no output array: ptar stands in for check_common's
    // output pointer, so only the input pointers and counts are checked
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, ptar)) return rc;
    if (ns > 0 && (!f || (kind == 1 && !n)))
        return fail(ctx, SE2P_ERR_ARG, "null density or normal pointer");
    return autotune(ctx, kind, psrc, ns, ptar, nt, f, n, Lx, Ly, prm, info);
}

int se2p_autotune_multibox(se2p_ctx *ctx, int kind, se2p_box_desc *boxes, int nbox, double tol,
                           se2p_tune_boxes_info *info)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (kind != 0 && kind != 1) return fail(ctx, SE2P_ERR_ARG, "kind must be 0 (SLP) or 1 (DLP)");
    return autotune_boxes(ctx, kind, boxes, nbox, tol, info);
}

int se2p_count_pairs(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                     int nside_x, int nside_y, double Lx, double Ly, int64_t *pairs_in,
                     int64_t *pairs_out)
{
    if (!ctx || !pairs_in || !pairs_out) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = refuse_in_group(ctx, "se2p_count_pairs")) return rc;
    double dummy_u = 0.0;
    if (int rc = check_common(ctx, psrc, ns, ptar, nt, nt > 0 ? &dummy_u : nullptr)) return rc;
    if (int rc = check_real(ctx, 1.0, nside_x, nside_y, Lx, Ly)) return rc;
    if (int rc = set_device(ctx)) return rc;
    *pairs_in = *pairs_out = 0;
    if (nt == 0) return SE2P_OK;
    HostIo io(ctx);
    io.want(psrc, 16 * (size_t)ns);
    io.want(ptar, 16 * (size_t)nt);
    if (int rc = io.map()) return rc;
    psrc = io.in(psrc);
    ptar = io.in(ptar);
    RealGeom g{psrc, ptar, ns, nt, 1.0, Lx, Ly, nside_x, nside_y};
    int64_t counts[2] = {0, 0};
    // densities are not read when counting; pass the positions as a stand-in
    if (int rc = real_sum(ctx, 1, g, psrc, psrc, 1, 2 * ns, 2 * nt, nullptr, counts)) return rc;
    *pairs_in = counts[0];
    *pairs_out = counts[1];
    return io.done(finish(ctx));
}

int se2p_selftest_keysort(se2p_ctx *ctx, const int *keys, int64_t n, int nkeys, int *perm,
                          int *off)
{
    if (!ctx) return SE2P_ERR_ARG;
    begin_call(ctx);
    if (int rc = refuse_in_group(ctx, "se2p_selftest_keysort")) return rc;
    if (n < 0 || n > 0x7fffffffLL || nkeys < 1 || (n > 0 && (!keys || !perm)) || !off)
        return fail(ctx, SE2P_ERR_ARG, "bad keysort arguments");
    if (int rc = set_device(ctx)) return rc;
    const size_t need = keysort_bytes(n, nkeys);
    if (int rc = ensure_workspace(ctx, need)) return rc;
    KeySort ks;
    keysort_carve(ctx->ws, n, nkeys, &ks);
    if (n > 0)
        SE2P_HIP(hipMemcpyAsync(ks.key, keys, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    if (int rc = keysort_run(ctx, &ks, n, nkeys)) return rc;
    if (n > 0)
        SE2P_HIP(hipMemcpyAsync(perm, ks.perm, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
    SE2P_HIP(hipMemcpyAsync(off, ks.off, sizeof(int) * (nkeys + 1), hipMemcpyDeviceToDevice,
                            ctx->stream));
    return finish(ctx);
}

double se2p_selftest_e1(double x)
{
    static double tab[SE2P_E1_NPIECE * SE2P_E1_ROW];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int p = 0; p < SE2P_E1_NPIECE; ++p) {
            double *row = tab + p * SE2P_E1_ROW;
            row[0] = se2p_e1_scale_shift[p][0];
            row[1] = se2p_e1_scale_shift[p][1];
            for (int k = 0; k < SE2P_E1_NCOEF; ++k) row[2 + k] = se2p_e1_coef[p][k];
            for (int k = 2 + SE2P_E1_NCOEF; k < SE2P_E1_ROW; ++k) row[k] = 0.0;
        }
    });
    if (!(x > 0)) return std::numeric_limits<double>::infinity();
    return se2p_e1(x, std::exp(-x), tab);
}

double se2p_selftest_exp_neg(double x)
{
    static const double tab[SE2P_EXP2_N] = SE2P_EXP2_TABLE;
    return se2p_exp_neg(x, tab);
}

void se2p_selftest_anchor(double x, double y, double Lx, double Ly, double h, int P, int *mx,
                          int *my, double *px, double *py)
{
    se2p_anchor(x, y, Lx, Ly, h, P, mx, my, px, py);
}

int se2p_selftest_cell(double x, double y, double Lx, double Ly, int nside_x, int nside_y)
{
    return se2p_cell_of(x, y, Lx, Ly, nside_x, nside_y);
}

}  // extern "C"
