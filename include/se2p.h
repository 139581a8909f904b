/*
 * se2p.h -- C ABI of the MI355X doubly periodic 2D Stokes spectral Ewald
 * library (libse2p.so).
 *
 * The four forward entry points replace the reference's four MEX kernels one
 * for one, with the same argument order and meaning:
 *
 *   se2p_slp_real    <- mex_stokes_slp_real(psrc,ptar,f,xi,nside_x,nside_y,Lx,Ly)
 *                       PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:58-83
 *   se2p_slp_kspace  <- mex_stokes_slp_kspace(psrc,ptar,xi,eta,f,Mx,My,Lx,Ly,w,P)
 *                       PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:8-42
 *   se2p_dlp_real    <- mex_stokes_dlp_real(psrc,ptar,f,n,xi,nside_x,nside_y,Lx,Ly)
 *                       PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_real.cpp:15-52
 *   se2p_dlp_kspace  <- mex_stokes_dlp_kspace(psrc,ptar,xi,eta,f,n,Mx,My,Lx,Ly,w,P)
 *                       PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_kspace.cpp:10-53
 *
 * Conventions
 *  - All arrays are FP64 DEVICE pointers (HBM), in the reference's layout: a
 *    point set, density or normal field of N entries is the MATLAB 2xN
 *    column-major matrix, i.e. interleaved [x0, y0, x1, y1, ...].
 *  - Outputs are caller-owned and fully overwritten (2 x Nt doubles).
 *  - Every call returns SE2P_OK (0) or an SE2P_ERR_* code; the message is
 *    available from se2p_last_error(ctx).  The reference aborts through
 *    mexErrMsgTxt where this returns a code.
 *  - A context owns a HIP stream, a device workspace and rocFFT plans.  One
 *    context per host thread.  Calls are synchronous (they return after the
 *    device work has finished) unless se2p_set_async(ctx, 1) was called, in
 *    which case they only enqueue work on the context's stream.
 *  - Integer parameters the reference reads as doubles and truncates
 *    (nside_x, nside_y, Mx, My, P) are passed as int.
 */
#ifndef SE2P_H
#define SE2P_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct se2p_ctx se2p_ctx;

/* status codes */
#define SE2P_OK 0
#define SE2P_ERR_ARG 1     /* null pointer, negative count, non-finite scalar */
#define SE2P_ERR_GRID 2    /* Mx/My odd, not >= P/2+2, or Lx/Mx != Ly/My */
#define SE2P_ERR_P 3       /* P odd or negative */
#define SE2P_ERR_W 4       /* w != P*Lx/(2*Mx) */
#define SE2P_ERR_NSIDE 5   /* nside_x or nside_y < 1 */
#define SE2P_ERR_PARAM 6   /* xi <= 0, eta <= 0, Lx or Ly <= 0 */
#define SE2P_ERR_NOCONV 7  /* parameter Newton iteration did not converge */
#define SE2P_ERR_HIP 10    /* HIP runtime error */
#define SE2P_ERR_FFT 11    /* rocFFT error */
#define SE2P_ERR_FROZEN 12 /* the call would allocate or free on a frozen context */

/* ---- context ----------------------------------------------------------- */
/* device >= 0: a HIP device.  device == SE2P_HOST_ONLY: a context that only
 * validates arguments (every call that would launch work returns
 * SE2P_ERR_HIP); for hosts without a GPU. */
#define SE2P_HOST_ONLY (-1)
int se2p_create(se2p_ctx **ctx, int device);
int se2p_destroy(se2p_ctx *ctx);
const char *se2p_last_error(const se2p_ctx *ctx);
/* Use a caller-provided hipStream_t (passed as void*); NULL restores the
 * context's own stream.  Work enqueued afterwards waits (by an event, without
 * blocking the host) for the work already enqueued on the previous stream,
 * since both use the context's workspace; while either stream is being
 * captured into a graph the caller orders them instead. */
int se2p_set_stream(se2p_ctx *ctx, void *hip_stream);
void *se2p_get_stream(se2p_ctx *ctx);
int se2p_set_async(se2p_ctx *ctx, int async_flag);
/* Host-pointer mode: with enable != 0, every array argument of the compute
 * entry points (points, densities, normals, du, outputs; for the multi-box
 * entries the arrays the descriptors point to) is host memory.  Inputs are
 * copied into a device staging area owned by the context, the computation
 * runs on the device, and outputs are copied back before the call returns
 * (such calls are synchronous whatever se2p_set_async says).  A host array
 * passed twice, such as targets that are the sources, is copied once.  The
 * default (0) is device pointers. */
int se2p_set_host_pointers(se2p_ctx *ctx, int enable);
/* Sorting of the cell lists and gridding bins.  Point sets of up to 229376
 * points with up to 14336 keys are sorted by a fused kernel sequence that
 * also gathers the sorted data (mode 0, the default); larger ones by a
 * multi-pass radix sort.  SE2P_SORT_RADIX forces the radix sort and
 * SE2P_SORT_BOX_KERNEL the separate run-bounding-box kernel, for validation:
 * every mode gives the same, bitwise identical results. */
#define SE2P_SORT_RADIX 1
#define SE2P_SORT_BOX_KERNEL 2
int se2p_set_sort_mode(se2p_ctx *ctx, int mode);
/* Passes of the Fourier-space backward call when the targets are the sources
 * and both sides' gradients are wanted without kept grids: 1 (the default)
 * runs both passes as one (the densities and du spread together, one
 * transform pair with both filters, one interpolation); 2 runs the target
 * pass and the source pass one after the other.  The results agree to
 * rounding.  For validation and measurement.  No reference counterpart. */
int se2p_set_adjoint_passes(se2p_ctx *ctx, int passes);
/* Forward state kept for the backward call.  With enable != 0, a
 * single-density forward call whose targets are its sources (ptar == psrc,
 * Nt == Ns; se2p_slp_real, se2p_dlp_real, se2p_slp_kspace, se2p_dlp_kspace)
 * keeps its sorted state in context-owned device memory: the real-space
 * cell lists, source records, chunk plans and run boxes; the Fourier-space
 * gridding bins, Gaussian factors, work plans and filtered grids.  The
 * matching backward call (same entry kind, same pointers, sizes and
 * parameters) then reuses it instead of sorting and, in Fourier space,
 * spreading and transforming again.  The caller must leave the arrays of
 * the forward call unchanged until the backward call, as for any saved
 * state; the library cannot detect an in-place change.  A later forward
 * call of the same kind replaces the state.  Host-pointer mode keeps
 * nothing.  Disabling frees the state.  Default 0.  No reference
 * counterpart. */
int se2p_set_keep(se2p_ctx *ctx, int enable);
/* Spreading work split: candidate sources per spreading workgroup (a tile's
 * candidate list is cut into chunks of this many, whose partial tiles are
 * added in chunk order).  0 selects the default (256 for one density, 128
 * for batched densities); otherwise a multiple of 16 in
 * [SE2P_SPREAD_CHUNK_MIN, SE2P_SPREAD_CHUNK_MAX].  The grids' summation order
 * follows the split, so results for different values differ by rounding;
 * for one value they are bitwise reproducible.  A change drops kept
 * Fourier-space state.  se2p_autotune sets it.  No reference counterpart. */
#define SE2P_SPREAD_CHUNK_MIN 64
#define SE2P_SPREAD_CHUNK_MAX 2048
int se2p_set_spread_chunk(se2p_ctx *ctx, int chunk);
/* Graph capture.  In async mode a call that finds its workspace, rocFFT
 * plans and kept state already sized neither allocates nor synchronises, so
 * it can be captured in a HIP graph.  A captured graph holds the device
 * addresses of the context's workspace, stream-slot workspaces, kept state
 * and rocFFT work buffers: it is valid only while those stay where they
 * are.  A later call that needs a larger workspace, a new grid size (new
 * plans), se2p_set_keep(ctx, 0), a change of se2p_set_spread_chunk and
 * se2p_destroy all move or free them.  With enable != 0 the context is
 * frozen: every call or setting that would allocate or free device memory
 * fails with SE2P_ERR_FROZEN instead (the outputs of such a call are
 * undefined; nothing is freed).  Freeze after
 * the warm-up calls that size the context and while graphs captured from it
 * exist; se2p_destroy is not guarded.  Default 0.  No reference
 * counterpart. */
int se2p_set_frozen(se2p_ctx *ctx, int enable);
/* Per-stage device timing with HIP events on the context stream.  When
 * enabled, se2p_stage_times fills up to n floats (milliseconds) for the
 * most recent call: sort, real, spread, fft_fwd, filter, fft_inv, interp,
 * pair (-1 for a stage the call did not run).  `real` is the whole
 * real-space stage; `pair` is its pair-sum kernel alone (without the
 * chunk planning and split-chunk sums around it). */
int se2p_set_timing(se2p_ctx *ctx, int enable);
int se2p_stage_times(se2p_ctx *ctx, float *ms, int n);
/* Concurrent groups.  The reference's drivers call the real-space and the
 * Fourier-space sum one after the other (StokesSLP_ewald_2p.m:119,126;
 * StokesDLP_ewald_2p.m:140,147), although neither reads the other's result.
 * Between se2p_group_begin and se2p_group_end, each call of a forward,
 * batched or backward entry point (the twelve se2p_{slp,dlp}_{real,kspace}
 * entries and their _batched and _backward forms) runs on a stream slot of
 * its own, with its own workspace, sort state and FFT plans, and the calls
 * of a group run side by side on the GPU.  Each waits for the work the
 * context stream held at se2p_group_begin; se2p_group_end makes the context
 * stream wait for all of them and, unless the context is in async mode,
 * waits for them on the host.  Outputs are defined only after
 * se2p_group_end, and are bitwise equal to the same calls made outside a
 * group.  The calls of one group must be independent: no call may write an
 * array another call of the group reads or writes.  At most
 * SE2P_GROUP_MAX_CALLS calls per group.  Inside a group, host-pointer mode,
 * the multi-box, driver, autotuning, pair-count and no-allocation entries
 * and the settings that bind a stream or free kept state fail with
 * SE2P_ERR_ARG, and stage times are not recorded.  A kept forward state
 * (se2p_set_keep) made in a group serves a backward call after
 * se2p_group_end.  No reference counterpart. */
#define SE2P_GROUP_MAX_CALLS 4
int se2p_group_begin(se2p_ctx *ctx);
int se2p_group_end(se2p_ctx *ctx);

/* ---- forward entry points (MEX replacements) ---------------------------- */
int se2p_slp_real(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                  int64_t nt, const double *f, double xi, int nside_x, int nside_y,
                  double Lx, double Ly, double *u);

int se2p_slp_kspace(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                    int64_t nt, double xi, double eta, const double *f, int Mx, int My,
                    double Lx, double Ly, double w, int P, double *u);

int se2p_dlp_real(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                  int64_t nt, const double *f, const double *n, double xi, int nside_x,
                  int nside_y, double Lx, double Ly, double *u);

int se2p_dlp_kspace(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                    int64_t nt, double xi, double eta, const double *f, const double *n,
                    int Mx, int My, double Lx, double Ly, double w, int P, double *u);

/* ---- the stresslet real-space sum without allocation ----------------------- */
/* No reference counterpart.  For callers that cannot allocate device memory
 * during a call: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same arguments, kernels and (bitwise) results as
 * se2p_dlp_real, but the call's sorted state (cell lists, source records,
 * chunk plans, run boxes) lives in a fixed device array of
 * SE2P_NOALLOC_SCRATCH_BYTES that the library holds on each GPU from its
 * first se2p_create on that GPU until the process ends.  The call makes no
 * hipMalloc, hipFree or other allocating call; the caller passes inputs and
 * the output as device pointers.
 *  - Limit: about 100 bytes per point (targets that are the sources; more
 *    with separate targets) plus 36 bytes per (cell, 4x4 sub-cell) key, so
 *    about 3.3e5 points on 22 x 22 cells.  Larger problems return
 *    SE2P_ERR_ARG with the bytes needed in se2p_last_error; use
 *    se2p_dlp_real for them.
 *  - Host-pointer mode (se2p_set_host_pointers) is rejected with
 *    SE2P_ERR_ARG: it stages arrays in allocated memory.
 *  - Every call on one GPU, from any context, uses that GPU's one scratch
 *    array, which keeps the last call's contents: calls that overlap in time
 *    on one GPU (two threads, or two streams not ordered by the caller) race
 *    and give wrong results.  Calls ordered on one stream are safe.
 *  - Nothing is kept for a backward call (se2p_set_keep does not apply). */
#define SE2P_NOALLOC_SCRATCH_BYTES (32u * 1024u * 1024u)
int se2p_dlp_real_noalloc(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, const double *n, double xi, int nside_x,
                          int nside_y, double Lx, double Ly, double *u);

/* ---- batched right-hand sides -------------------------------------------- */
/* nrhs density vectors on one geometry (same sources, targets and, for the
 * DLP, normals).  Density r is at f + r*f_stride and result r at
 * u + r*u_stride (strides in doubles, f_stride >= 2*Ns, u_stride >= 2*Nt).
 * The cell lists, gridding bins, Gaussian factors and every pair kernel
 * evaluation are shared by the right-hand sides, and the Fourier-space sum
 * runs one batched transform pair for all of them.  Result r equals the
 * single-density entry point's result for density r (to rounding).  The
 * reference has no batched form; its callers loop over the MEX kernels. */
int se2p_slp_real_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, int nrhs, int64_t f_stride, double xi,
                          int nside_x, int nside_y, double Lx, double Ly, double *u,
                          int64_t u_stride);
int se2p_slp_kspace_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                            int64_t nt, double xi, double eta, const double *f, int nrhs,
                            int64_t f_stride, int Mx, int My, double Lx, double Ly, double w, int P,
                            double *u, int64_t u_stride);
int se2p_dlp_real_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                          int64_t nt, const double *f, const double *n, int nrhs, int64_t f_stride,
                          double xi, int nside_x, int nside_y, double Lx, double Ly, double *u,
                          int64_t u_stride);
int se2p_dlp_kspace_batched(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                            int64_t nt, double xi, double eta, const double *f, const double *n,
                            int nrhs, int64_t f_stride, int Mx, int My, double Lx, double Ly,
                            double w, int P, double *u, int64_t u_stride);

/* ---- many independent boxes in one call ------------------------------------ */
/* One periodic problem per box: its own sources, targets, densities, box size
 * and Ewald parameters (for instance many periodic cells of a simulation).
 * Result b equals the single-box entry point's result on box b (the same
 * kernels; to rounding).  All pointers are device arrays.  The real-space
 * sums of all boxes run as one sort and one pair-kernel launch; in Fourier
 * space every box's sorts and transforms run on four stream slots and the
 * spreading and interpolation of up to 16 boxes share a launch.  The
 * reference has no multi-box form; its callers loop over boxes. */
typedef struct se2p_box_desc {
    const double *psrc; int64_t ns;   /* Ns x 2 sources */
    const double *ptar; int64_t nt;   /* Nt x 2 targets */
    const double *f;                  /* Ns x 2 densities */
    const double *n;                  /* Ns x 2 normals (stresslet; NULL for the Stokeslet) */
    double Lx, Ly;
    double xi;                        /* Ewald parameter */
    int nside_x, nside_y;             /* real-space cells */
    double eta, w;                    /* Fourier space: eta, Gaussian support half-width */
    int Mx, My, P;                    /* Fourier space: grid and support */
    double *u;                        /* Nt x 2 result (overwritten) */
} se2p_box_desc;

int se2p_slp_real_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox);
int se2p_slp_kspace_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox);
int se2p_dlp_real_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox);
int se2p_dlp_kspace_multibox(se2p_ctx *ctx, const se2p_box_desc *boxes, int nbox);

/* ---- adjoints (backward entry points) ------------------------------------ */
/* Gradients of L = sum_t du_t . u_t, u the output of the matching forward
 * entry point on the same arguments (one density): dL/dpsrc and dL/df (Ns x 2),
 * dL/dptar (Nt x 2) and, for the stresslet, dL/dn (Ns x 2).  du is Nt x 2.
 * Any output pointer may be NULL; that gradient is then not computed.
 * Treated as fixed (piecewise constant in the positions): the cell of every
 * point, the cutoff decision of every pair and the grid anchor of every point.
 * A coincident pair (r^2 == 0) enters the forward sum as a constant (the
 * Stokeslet self term) or not at all (stresslet) and contributes no position
 * gradient; dL/dpsrc + dL/dptar is then the gradient for sources and targets
 * moved together (targets == sources).  The reference has no adjoint: these
 * are the transposes of mex_stokes_slp_real.cpp, mex_stokes_slp_kspace.cpp,
 * mex_stokes_dlp_real.cpp and mex_stokes_dlp_kspace.cpp as restated above. */
int se2p_slp_real_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                           int64_t nt, const double *f, double xi, int nside_x, int nside_y,
                           double Lx, double Ly, const double *du, double *dpsrc, double *dptar,
                           double *df);
int se2p_slp_kspace_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                             int64_t nt, double xi, double eta, const double *f, int Mx, int My,
                             double Lx, double Ly, double w, int P, const double *du,
                             double *dpsrc, double *dptar, double *df);
int se2p_dlp_real_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                           int64_t nt, const double *f, const double *n, double xi, int nside_x,
                           int nside_y, double Lx, double Ly, const double *du, double *dpsrc,
                           double *dptar, double *df, double *dn);
int se2p_dlp_kspace_backward(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                             int64_t nt, double xi, double eta, const double *f, const double *n,
                             int Mx, int My, double Lx, double Ly, double w, int P,
                             const double *du, double *dpsrc, double *dptar, double *df,
                             double *dn);

/* ---- the reference drivers as single calls --------------------------------- */
/* Parameters of the spectral Ewald split.  Inputs left at 0 take the
 * reference driver's defaults: P = 24, Nb = min(P*round(log2(Ns+Nt)), (Ns+Nt)/4),
 * tol = 1e-16, grid M = min(2*kinf, 10000) (Mx = A*M, My = B*M with
 * [A,B] = rat(Lx/Ly)).  The other fields are outputs.
 *
 * fixed_grid != 0 (input; no reference counterpart) with M > 0 given: the
 * grid is fixed and the drivers' estimates are inverted the other way round:
 * xi so that the k-space estimate at kinf = M/2 equals tol
 * (StokesSLP_ewald_2p.m:178-203, StokesDLP_ewald_2p.m:200-225), then the
 * smallest cutoff whose real-space estimate at that xi is tol (:147-176,
 * :171-198), then the largest nside_x = a*A, nside_y = a*B whose cells are at
 * least that cutoff wide.  Nb is ignored and returned as the resulting points
 * per cell; rc is the cell side. */
typedef struct se2p_params {
    int P;
    double Nb;
    double tol;
    int M;
    int nside_x, nside_y, Mx, My, kinf;
    int converged;   /* 0 if a Newton solve hit its iteration cap (the reference warns) */
    double rc, xi, eta, w;
    int fixed_grid;  /* input: 1 selects the fixed-grid mode above (needs M > 0) */
} se2p_params;

/* Parameter selection of StokesSLP_ewald_2p.m:27-104 (kind 0) or
 * StokesDLP_ewald_2p.m:28-125 (kind 1); Q = sum(f.^2) + 1.  Host only. */
int se2p_select_params(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                       se2p_params *p);

/* [u1,u2,ur,uk,xi] = StokesSLP_ewald_2p(xsrc,ysrc,xtar,ytar,f1,f2,Lx,Ly,...)
 * (PeriodicEwald/src/matlab/StokesSLP_ewald_2p.m:1-138): parameters from
 * *prm (filled on return), points wrapped into [-L/2, L/2), real + Fourier
 * sums, u = ur + uk.  psrc/ptar are the caller's (unwrapped) points; ur and
 * uk may be NULL.  All arrays are device arrays.  The two sums run side by
 * side on the GPU, as the two calls of a concurrent group would (the
 * Fourier-space sum on the context stream, the real-space sum on a stream
 * slot), and give bitwise the results of the same sums run one after the
 * other. */
int se2p_slp_ewald(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                   const double *f, double Lx, double Ly, se2p_params *prm, double *u, double *ur,
                   double *uk);
/* [u1,u2,ur,uk,xi] = StokesDLP_ewald_2p(xsrc,ysrc,xtar,ytar,n1,n2,f1,f2,Lx,Ly,...)
 * (StokesDLP_ewald_2p.m:1-163): as above, then uk -= sum((f.n) x_src)/(Lx Ly)
 * with the caller's unwrapped source points, and u = -(ur + uk).  The
 * driver's renormalisation of the normals on MATLAB releases after R2017a
 * (:69-80) is not applied: normals are used as given. */
int se2p_dlp_ewald(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                   const double *f, const double *n, double Lx, double Ly, se2p_params *prm,
                   double *u, double *ur, double *uk);

/* ---- parameter autotuning (no reference counterpart) --------------------- */
#define SE2P_TUNE_MAX 16
#define SE2P_TUNE_CHUNKS 3
#define SE2P_TUNE_PCAND 6
typedef struct se2p_tune_info {
    int ncand;                 /* candidates timed (0 on a cache hit) */
    int best;                  /* index of the chosen candidate */
    int cached;                /* 1: the choice came from the per-process cache */
    int P;                     /* the Gaussian support used (prm->P, or the drivers' 24) */
    double Nb[SE2P_TUNE_MAX];  /* candidate points per cell */
    int nside_x[SE2P_TUNE_MAX];
    int Mx[SE2P_TUNE_MAX];
    double ms[SE2P_TUNE_MAX];  /* measured real + Fourier step of each candidate, ms */
    int spread_chunk;          /* chosen spreading work split (0: the default) */
    int chunk[SE2P_TUNE_CHUNKS];        /* spreading splits timed at the chosen Nb */
    double chunk_ms[SE2P_TUNE_CHUNKS];  /* their real + Fourier step, ms */
    int npc;                            /* support/grid candidates timed (stage 3) */
    int pc_P[SE2P_TUNE_PCAND];          /* their support P */
    int pc_Mx[SE2P_TUNE_PCAND];         /* their grid */
    double pc_ms[SE2P_TUNE_PCAND];      /* their real + Fourier step, ms */
    double pc_dev[SE2P_TUNE_PCAND];     /* max |uk - uk(P = 24)| / max |ur + uk(P = 24)|;
                                           -1: not checked (not faster) */
} se2p_tune_info;

/* Chooses the drivers' free parameters for one geometry at prm->tol
 * (required, 0 < tol < 1) by timing the real- plus Fourier-space sums on
 * this context, side by side as the driver entries run them (each
 * candidate: a warm-up step, then the best of three blocks of three
 * back-to-back steps).  Three stages:
 *  1. the points per cell Nb, which sets the cell grid (nside), the
 *     splitting xi and the Fourier grid through the driver formulas
 *     (StokesSLP_ewald_2p.m:33-104, StokesDLP_ewald_2p.m:33-125):
 *     Nb = Nb0 * 2^(k/2), k = -6..2, around the drivers' default
 *     Nb0 = min(P*round(log2(Ns+Nt)), (Ns+Nt)/4), one candidate per cell
 *     grid, grids above 4096^2 points skipped.  Every candidate meets the
 *     drivers' truncation estimates at tol.
 *  2. at that Nb, the spreading work split (se2p_set_spread_chunk): the
 *     default, 128 and 512; the fastest is set on the context.
 *  3. only when prm->P is 0: supports P = 16 and 20 on grids 1, 1.2 and 1.4
 *     times the drivers' grid (the grid override M).  A candidate faster
 *     than the best so far is taken only if its Fourier-space sum differs
 *     from the P = 24 one by at most tol/4 of max |ur + uk| (P = 24): the
 *     Fourier-space sum converges as the support and the grid grow, so that
 *     difference bounds the error the candidate adds.
 * P is prm->P, or the drivers' 24 when 0.  prm->M must be 0 and
 * prm->fixed_grid 0.  On return *prm holds the chosen parameters as
 * se2p_select_params fills them for its Nb, P and M: passing *prm on to
 * se2p_slp_ewald / se2p_dlp_ewald reproduces them.  Choices are cached per
 * process by (device, kind, Ns, Nt, targets-are-sources, Lx, Ly, tol, P
 * requested); a cache hit times nothing and sets the cached spreading split
 * on the context.  Points are the caller's (unwrapped) points, as for the
 * driver entries.  Synchronous.  info may be NULL. */
int se2p_autotune(se2p_ctx *ctx, int kind, const double *psrc, int64_t ns, const double *ptar,
                  int64_t nt, const double *f, const double *n, double Lx, double Ly,
                  se2p_params *prm, se2p_tune_info *info);

/* Multi-box tuning (no reference counterpart).  Writes each box's Ewald
 * parameters (xi, nside_x/_y, eta, w, Mx, My, P) into its descriptor: the
 * drivers' parameters at tol with P = 24 (se2p_select_params, Nb default)
 * timed as one multi-box step (real- plus Fourier-space multi-box calls),
 * then each box's points per cell tuned on its own as stage 1 of
 * se2p_autotune (or taken from se2p_autotune's per-process cache when that
 * geometry was tuned there with P = 24 at the same tol), and the spreading work split (default, 128, 512) of the
 * multi-box step at those parameters.  The tuned set is kept if its best
 * step is faster than the drivers' set, otherwise the drivers' parameters
 * and the default split are.  Last, every grid is rounded up to the next
 * even length 2^a 3^b 5^c 7^d (the grid override M; xi, the cells and P
 * unchanged, so the Fourier-space truncation error can only fall), kept if
 * the step is faster.  The split is set on the context.  Every box
 * needs at least two points; points are wrapped, as for the multi-box
 * entries; u is overwritten.  Device arrays only.  Synchronous; info may be
 * NULL. */
typedef struct se2p_tune_boxes_info {
    double drivers_ms;                  /* multi-box step at the drivers' parameters */
    double tuned_ms;                    /* best step at the parameters written */
    int kept_drivers;                   /* 1: the drivers' parameters were kept */
    int spread_chunk;                   /* the split set on the context */
    int chunk[SE2P_TUNE_CHUNKS];        /* splits timed at the tuned cells */
    double chunk_ms[SE2P_TUNE_CHUNKS];  /* their multi-box step, ms */
    double smooth_ms;                   /* step with the grids rounded up to 2^a 3^b 5^c 7^d
                                           (-1: every grid already was) */
    int smooth_grids;                   /* grids rounded up in the parameters written (0: none) */
} se2p_tune_boxes_info;

int se2p_autotune_multibox(se2p_ctx *ctx, int kind, se2p_box_desc *boxes, int nbox, double tol,
                           se2p_tune_boxes_info *info);

/* ---- pair statistics (no reference counterpart) ------------------------- */
/* Number of source-target pairs the real-space sum evaluates (pairs_in: same
 * cell, or neighbour cell with r^2 < cutoff^2) and rejects (pairs_out).  Used
 * to state the real-space sum's algorithmic flops. */
int se2p_count_pairs(se2p_ctx *ctx, const double *psrc, int64_t ns, const double *ptar,
                     int64_t nt, int nside_x, int nside_y, double Lx, double Ly,
                     int64_t *pairs_in, int64_t *pairs_out);

/* ---- self-test hooks (no reference counterpart) -------------------------- */
/* Host builds of the device functions the kernels use, so that a CPU-only test
 * suite can pin their arithmetic: the E1 approximation, the gridding anchor
 * (same arithmetic as FindClosestNode, ewald_tools.cpp:86-114) and the cell
 * index (ewald_tools.cpp:28-35). */
double se2p_selftest_e1(double x);
/* exp(-x) as the real-space kernels evaluate it (x >= 0). */
double se2p_selftest_exp_neg(double x);
void se2p_selftest_anchor(double x, double y, double Lx, double Ly, double h, int P, int *mx,
                          int *my, double *px, double *py);
int se2p_selftest_cell(double x, double y, double Lx, double Ly, int nside_x, int nside_y);
/* The device sort behind the cell lists and gridding bins: stable order of
 * keys[0..n) (device array, values in [0, nkeys)) into perm[0..n) and run
 * offsets off[0..nkeys] (device arrays). */
int se2p_selftest_keysort(se2p_ctx *ctx, const int *keys, int64_t n, int nkeys, int *perm,
                          int *off);

#ifdef __cplusplus
}
#endif

#endif /* SE2P_H */
