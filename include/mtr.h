/*
 * mtr.h -- C ABI of the MI355X multi-axis tomography reconstructor.
 *
 * Plain C: every argument is a POD struct, a pointer or a size; no HIP,
 * MPI or torch types cross this boundary.  Every function returns an int
 * status (MTR_OK == 0, a negative MTR_ERR_* otherwise) and never exits the
 * process; mtr_last_error() describes the last failure of the calling
 * thread.
 *
 * The entry points replace the reference's GPU host wrappers and the parts
 * of its driver that form the SIRT / MACE hot path (file:line in
 * Autumn1998/Multi-tilt-Reconstuction):
 *   mtr_bpt            ComputeInitialModel      FunctionOnGPU.cu:138-179 (declared FunctionOnGPU.cuh:41)
 *   mtr_sirt           runSIRTOnGPU             FunctionOnGPU.cu:182-266 (declared FunctionOnGPU.cuh:42)
 *   mtr_form_z         initialZdata             FunctionOnGPU.cu:291-313 (declared FunctionOnGPU.cuh:43)
 *   mtr_mann           runMaceOnCPU             MultiAxis.cpp:33-38 (GPU twin runMaceOnGPU FunctionOnGPU.cuh:44)
 *   mtr_op_form_z      initialZOnGPU            FunctionOnGPU.cu:268-276
 *   mtr_op_mann        runMaceOnCPU             MultiAxis.cpp:33-38
 *   mtr_op_mean        consensus average        MultiAxis.cpp:64-65
 *   mtr_mean           consensus average        MultiAxis.cpp:61-66
 *   mtr_mace_init      BPT-sum initialisation   MultiAxis.cpp:163-200
 *   mtr_mace_solve     reconstructModel         MultiAxis.cpp:40-71 (+ the initialisation above)
 *   mtr_op_fp          computeDivisor + Reproj  FunctionOnGPU.cu:316-333, :27-52
 *   mtr_op_residual    host residual loop       FunctionOnGPU.cu:228-236
 *   mtr_op_bp          sirtBackProjOnGPU        FunctionOnGPU.cu:111-135 (+ BilinearValue :54-79)
 *   mtr_op_bpt         backProjOnGPU            FunctionOnGPU.cu:82-109
 *   mtr_divide_slice   divideSlice              MultiAxis.cpp:13-31
 * The volume and projection descriptors are layout-identical to the
 * reference's VolumeSize and ProjectionSize (MrcFileIOBaseFunction.h:57-78);
 * geometry is the reference's per-view affine coefficients
 * xc[4*AngN], yc[4*AngN] (u = xc0 + xc1*X + xc2*Y + xc3*Z - 1,
 * FunctionOnGPU.cu:12-25).
 *
 * Threading: one context per device and host thread; a context is not
 * internally locked.  Device state (projections, geometry, volumes) stays
 * resident in HBM for the context's lifetime; host output pointers are
 * optional so a device-only pipeline never crosses PCIe.
 */
#ifndef MTR_H_
#define MTR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTR_ABI_VERSION 1

enum {
    MTR_OK = 0,
    MTR_ERR_ARG = -1,       /* null pointer / bad enum / bad count            */
    MTR_ERR_SHAPE = -2,     /* inconsistent or unsupported sizes              */
    MTR_ERR_OOM = -3,       /* device allocation failed / does not fit        */
    MTR_ERR_HIP = -4,       /* HIP runtime error                              */
    MTR_ERR_NONFINITE = -5, /* NaN or Inf in an input                         */
    MTR_ERR_STATE = -6,     /* call order (e.g. solve before agents)          */
    MTR_ERR_RANGE = -7,     /* index out of range                             */
    MTR_ERR_NODEVICE = -8   /* no usable GPU                                  */
};

/* Layout-identical to VolumeSize, MrcFileIOBaseFunction.h:57-70. */
typedef struct {
    int Xstart, Xend, Ystart, Yend, Zstart, Zend;
    int X, Y, Z;
} mtr_volume_size;

/* Layout-identical to ProjectionSize, MrcFileIOBaseFunction.h:73-78
 * (X = detector columns nu, Y = detector rows nv, AngN = views). */
typedef struct {
    int X, Y, AngN;
} mtr_projection_size;

/* Layout-identical to Slice, MrcFileIOBaseFunction.h:50-54. */
typedef struct {
    int z_start, z_end;
} mtr_slice;

/* Storage precision of volumes and projections.  Ray and back-projection
 * sums always accumulate in FP32; residual and consensus norms in FP64. */
enum { MTR_PREC_FP32 = 0, MTR_PREC_BF16 = 1, MTR_PREC_FP16 = 2 };

/* Execution form of mtr_mace_solve / mtr_sirt. */
enum {
    MTR_FORM_FAST = 0,  /* fused, batched over agents, graph-replayed        */
    MTR_FORM_PLAIN = 1  /* one agent at a time, one launch per phase, host
                           sync per iteration, splat FP with float atomics
                           and the normaliser recomputed every iteration:
                           the reference's structure on this GPU           */
};

/* Volumes addressable by mtr_get_volume / mtr_set_volume. */
enum {
    MTR_VOL_WORK = 0, /* agent working volume (z before SIRT, v after)       */
    MTR_VOL_W = 1,    /* agent MACE state w_a                                */
    MTR_VOL_WBAR = 2  /* consensus w-bar (agent index ignored)               */
};

typedef struct mtr_ctx mtr_ctx;

int mtr_abi_version(void);
const char *mtr_last_error(void);

/* ---- context ------------------------------------------------------- */
int mtr_ctx_create(int device, mtr_ctx **out);
int mtr_ctx_destroy(mtr_ctx *ctx);
/* Storage precision for agents added afterwards (default FP32). */
int mtr_ctx_set_precision(mtr_ctx *ctx, int precision);
/* Fast solves capture one MACE step into a HIP graph and replay it for
 * every step (default on; off: every launch is issued from the host, as
 * also happens while per-kernel timing is enabled).  No counterpart in the
 * reference. */
int mtr_ctx_use_graphs(mtr_ctx *ctx, int on);
/* Back-projection tiles are chosen per agent by timing the candidates on
 * the agent's own data at its first BPT (mtr_bpt, mtr_mace_init); choices
 * are cached per shape and geometry, and in the file MTR_TUNE_CACHE names.
 * Off: fixed defaults.  No counterpart in the reference (fixed 8x8x8
 * blocks, ReadAndInitialOptions.h:29-31). */
int mtr_ctx_autotune(mtr_ctx *ctx, int on);
/* Tile choice of an agent: voxels per thread along the run axis (0 for
 * the general kernel) and the axis its 64 lanes run along (0 = X,
 * 1 = Y, 2 = Z). */
int mtr_agent_tiles(const mtr_ctx *ctx, int agent, int *run, int *lanes_axis);
/* Detector row slots each wave of the column-plan forward projection keeps
 * in LDS: a window that the strip walk writes out as it moves (a power of
 * two), or 0 for whole rows (and for agents without the column plan).
 * Diagnostic; no counterpart in the reference.  MTR_FP_COL_WINDOW=0 in the
 * environment when an agent is added keeps whole rows. */
int mtr_agent_fp_window(const mtr_ctx *ctx, int agent, int *slots);
/* The reconstruction box shared by every agent (the union box the
 * reference computes in synchronizeVol, MultiAxis.cpp:73-99).  Must be set
 * before the first agent is added; allocates w-bar. */
int mtr_ctx_set_volume(mtr_ctx *ctx, const mtr_volume_size *vol);

/* ---- agents (one tilt series each) ---------------------------------- */
/* Uploads projections p[AngN][Y][X] and geometry once; they stay resident.
 * Allocates the agent's working volume, MACE state, residual and
 * normaliser.  MTR_ERR_NONFINITE for a NaN or Inf in p or the geometry. */
int mtr_agent_add(mtr_ctx *ctx, const mtr_projection_size *prj, const double *xc,
                  const double *yc, const float *p_host, int *agent_id);
/* Same, with p already in device memory (not copied; must stay valid
 * until the context is destroyed).  p_dev is float for every precision
 * and is converted into the agent's own storage; it is not scanned for
 * non-finite values (the geometry is). */
int mtr_agent_add_device(mtr_ctx *ctx, const mtr_projection_size *prj, const double *xc,
                         const double *yc, const float *p_dev, int *agent_id);
int mtr_agent_count(const mtr_ctx *ctx, int *count);
/* Upper bound of the HBM bytes an agent of this shape occupies at the
 * context precision (whatever kernels its geometry selects). */
int mtr_agent_footprint(const mtr_ctx *ctx, const mtr_projection_size *prj, size_t *bytes);
/* Device bytes an added agent holds now, and the most it held (staging
 * buffers of the upload included); mtr_agent_footprint bounds both.  No
 * counterpart in the reference. */
int mtr_agent_device_bytes(const mtr_ctx *ctx, int agent, size_t *resident, size_t *peak);
/* Kernels the geometry of an agent selected (no counterpart in the
 * reference, which has one kernel per operator):
 *   forward projection  MTR_FPK_GENERAL  ray-driven gather over the planes
 *                                        of the driving axis
 *                       MTR_FPK_STREAM   row-separable views: one ray per
 *                                        wave over 64 adjacent planes
 *                       MTR_FPK_COLUMN   v fixed per (X, Y) column in every
 *                                        view: strip accumulation
 *   back-projection     MTR_BPK_GENERAL, MTR_BPK_SEPARABLE (runs along the
 *                                        plane axis), MTR_BPK_COLUMN (runs
 *                                        along Z) */
#define MTR_FPK_GENERAL 0
#define MTR_FPK_STREAM 1
#define MTR_FPK_COLUMN 2
#define MTR_BPK_GENERAL 0
#define MTR_BPK_SEPARABLE 1
#define MTR_BPK_COLUMN 2
int mtr_agent_kernels(const mtr_ctx *ctx, int agent, int *fp_kind, int *bp_kind);

/* ---- the reference's hot-path entry points, device-resident ---------- */
/* BPT of one agent into its working volume (zeroed first).  out_host
 * (nullable) receives the volume. */
int mtr_bpt(mtr_ctx *ctx, int agent, float *out_host);
/* iters SIRT iterations (the reference hard-codes 10) of one agent on its
 * working volume.  x0_host (nullable) overwrites the working volume first;
 * out_host (nullable) receives the result. */
int mtr_sirt(mtr_ctx *ctx, int agent, const float *x0_host, int iters, float lambda,
             float *out_host);
/* z_a = 2*wbar - w_a into the agent's working volume. */
int mtr_form_z(mtr_ctx *ctx, int agent);
/* w_a = rho*(2*v_a - z_a) + (1-rho)*w_a with v_a the working volume and
 * z_a = 2*wbar - w_a recomputed from the state the step started from. */
int mtr_mann(mtr_ctx *ctx, int agent, float rho);
/* wbar = mean over agents of w_a. */
int mtr_mean(mtr_ctx *ctx);
/* BPT of every agent, summed: w_a = wbar = sum_a BPT_a. */
int mtr_mace_init(mtr_ctx *ctx);
/* The whole reconstruction: mtr_mace_init then M MACE steps of
 * S SIRT iterations per agent.  wbar_host (nullable) receives w-bar. */
int mtr_mace_solve(mtr_ctx *ctx, int M, int S, float lambda, float rho, int form,
                   float *wbar_host);
/* Convergence control and checkpoint/resume (no counterpart in the
 * reference, which always runs -n steps and keeps no state between runs).
 * Continues a fast MACE solve from the current state w_a, w-bar: after
 * mtr_mace_init or a solve, or after mtr_set_volume of MTR_VOL_W for every
 * agent and of MTR_VOL_WBAR (resume from a checkpoint taken with
 * mtr_get_volume).  Runs at most M steps; with rel_tol > 0 it stops after
 * the first step whose FP64 consensus difference sum_a ||w_a - wbar||^2
 * changed by at most rel_tol times the previous step's value.  Continuing
 * is bit-identical to not having stopped.  steps_run and wbar_host are
 * optional. */
int mtr_mace_continue(mtr_ctx *ctx, int M, int S, float lambda, float rho, double rel_tol,
                      int *steps_run, float *wbar_host);

/* ---- state access ---------------------------------------------------- */
int mtr_get_volume(mtr_ctx *ctx, int which, int agent, float *host);
int mtr_set_volume(mtr_ctx *ctx, int which, int agent, const float *host);
/* MACE checkpoint files (no counterpart in the reference, which keeps no
 * state between runs): save writes the state w_0 .. w_{K-1}, w-bar as
 * float32 volumes after a 64-byte header (magic "MTRCKPT1", K, the box, the
 * caller's count of steps done, the storage precision, a hash of every
 * agent's detector size and coefficients; the projection values are not
 * hashed); save writes a temporary file beside the target and renames it
 * over the target once written and flushed, so an interrupted save leaves
 * an earlier checkpoint at that path intact; load checks magic, agent
 * count, box, geometry hash, storage precision and file length
 * (MTR_ERR_SHAPE), and NaN or Inf (MTR_ERR_NONFINITE) before it replaces the
 * state, then returns the step count.  mtr_mace_continue resumes from a
 * loaded state bit-exactly. */
int mtr_checkpoint_save(mtr_ctx *ctx, const char *path, int steps_done);
int mtr_checkpoint_load(mtr_ctx *ctx, const char *path, int *steps_done);
/* Device pointer of w-bar (FP32 storage only; for device-side consumers). */
int mtr_wbar_device(mtr_ctx *ctx, void **dev_ptr);

/* ---- convergence norms (FP64) ---------------------------------------- */
/* sum of r^2 over the last SIRT iteration's residual of an agent. */
int mtr_residual_norm2(mtr_ctx *ctx, int agent, double *out);
/* sum over agents of ||w_a - wbar||^2 after the last consensus step. */
int mtr_consensus_diff2(mtr_ctx *ctx, double *out);

/* ---- stateless operators (host buffers in, host buffers out) --------- */
/* Splat-equivalent forward projection of volume x (vol box, [Z][Y][X])
 * into s and c ([AngN][Y][X]), computed as a deterministic ray-driven
 * gather. */
int mtr_op_fp(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol,
              const double *xc, const double *yc, const float *x, float *s, float *c);
/* Allocation-free variant of mtr_op_fp, on device buffers: x_dev (vol box,
 * [Z][Y][X]) in, s_dev and c_dev ([AngN][Y][X]) out, all float device
 * memory of the context's GPU (MTR_ERR_ARG otherwise).  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc or hipFree: the view tables go to a
 * fixed per-GPU scratch array of the library (61 KB, held until the process
 * ends), so at most MTR_NOALLOC_MAX_VIEWS views per call (MTR_ERR_SHAPE
 * beyond).  Calls that overlap on one GPU (from several host threads or
 * contexts) share that array and race; calls in sequence do not.  Returns
 * after the projection has finished.  Same kernel and results as
 * mtr_op_fp. */
#define MTR_NOALLOC_MAX_VIEWS 256
int mtr_op_fp_noalloc(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol,
                      const double *xc, const double *yc, const float *x_dev, float *s_dev, float *c_dev);
/* r = p - (c != 0 ? s / c : s). */
int mtr_op_residual(mtr_ctx *ctx, int64_t n, const float *p, const float *s, const float *c,
                    float *r);
/* x += fl32(S/C) * lambda where C != 0 (in place on x). */
int mtr_op_bp(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol,
              const double *xc, const double *yc, const float *r, float *x, float lambda);
/* x = fl32(S/C) of p (x need not be initialised). */
int mtr_op_bpt(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol,
               const double *xc, const double *yc, const float *p, float *x);
/* z = 2*wbar - w (initialZOnGPU, FunctionOnGPU.cu:268-276). */
int mtr_op_form_z(mtr_ctx *ctx, int64_t n, const float *w, const float *wbar, float *z);
/* w = rho*(2v - z) + (1-rho)*w in place (runMaceOnCPU, MultiAxis.cpp:33-38). */
int mtr_op_mann(mtr_ctx *ctx, int64_t n, const float *v, float *w, const float *z, float rho);
/* wbar = (w[0] + ... + w[K-1]) / K (MultiAxis.cpp:64-65). */
int mtr_op_mean(mtr_ctx *ctx, int64_t n, int K, const float *const *w, float *wbar);
/* Z-slab partition of the reference (MultiAxis.cpp:13-31). */
int mtr_divide_slice(int myid, int nparts, int Z, int Zstart, mtr_slice *out);

/* ---- kernel timing (HIP events on the context's stream) -------------- */
enum {
    MTR_K_FP = 0,        /* forward projection + fused residual            */
    MTR_K_BP = 1,        /* back-projection update                          */
    MTR_K_BPT = 2,       /* initial back-projection                         */
    MTR_K_CONSENSUS = 3, /* Mann + mean + next z (fused)                    */
    MTR_K_NORMALISER = 4,/* geometry-only FP normaliser                     */
    MTR_K_ELEMENTWISE = 5,/* standalone z / Mann / mean / residual          */
    MTR_K_SPLAT = 6,     /* plain form: atomic splat FP                     */
    /* Per-kernel parts of the classes above, counted in addition to them
     * (SIRT iterations only, not the initial BPT):                        */
    MTR_K_FP_STREAM = 7,     /* row-separable streaming FP + residual      */
    MTR_K_FP_COLUMN = 8,     /* column-plan FP into partial rows           */
    MTR_K_FP_COLUMN_RES = 9, /* column-plan row sums + residual            */
    MTR_K_FP_GENERAL = 10,   /* general-plan FP + residual                 */
    MTR_K_BP_SEPARABLE = 11, /* runs along the plane axis                  */
    MTR_K_BP_COLUMN = 12,    /* runs along Z                               */
    MTR_K_BP_GENERAL = 13,   /* general plan                               */
    MTR_K_FP_STREAM_RES = 14,/* Z-segmented streaming FP: sums + residual  */
    MTR_K_COUNT = 15
};
/* Enable (1) or disable (0) per-launch event timing.  While it is on, the
 * SIRT iterations keep every kernel on the context's stream (by default
 * the column-plan agents and the others run side by side on two), so each
 * kernel is timed alone; results are the same either way. */
int mtr_timing_enable(mtr_ctx *ctx, int on);
int mtr_timing_reset(mtr_ctx *ctx);
/* Launch count and summed event time (ms) of one kernel class. */
int mtr_timing_get(mtr_ctx *ctx, int kernel, int64_t *launches, double *ms);
/* Copies `bytes` between two fresh device buffers with the library's
 * streaming copy kernel and returns the kernel time; used to calibrate HBM
 * traffic counters against a copy of known size. */
int mtr_calibrate_copy(mtr_ctx *ctx, size_t bytes, double *ms);
/* Blocks until all queued work of the context is done. */
int mtr_sync(mtr_ctx *ctx);

/* ---- request server (continuous batching) ----------------------------
 * No counterpart in the reference, which runs one reconstruction per MPI
 * job (MultiAxis.cpp:101-230).  A server owns one device stream, four
 * preparation workers (threads, each with a stream of its own) and a
 * retirer thread with a stream of its own.  A
 * request is a volume box and K tilt series (each with its own detector
 * size and geometry) plus its solve parameters (M MACE steps of S SIRT
 * iterations, lambda, rho); the server copies it on submission.  Requests
 * are admitted (in arrival order unless mtr_server_set_admission says
 * otherwise) while their HBM footprint fits the budget and fewer than
 * max_active are active.  The workers prepare the requests the policy
 * takes next (device memory, geometry tables, uploads; their bytes count
 * as in use from then on) while ticks run, and admission forms the
 * BPT-sum initial model.  Server calls are for one host thread.  Each
 * mtr_server_step first completes the previous step (waits for its device work; the requests
 * that finished with it count as done and go to a second thread, the
 * retirer, which copies their w-bar to the host and frees their device
 * memory while the next steps run), then admits, then launches one MACE
 * step of every active request, the SIRT iterations of all their agents
 * batched into shared launches, and returns while that step runs, so the
 * caller's work between steps overlaps the GPU's.  A request whose last
 * MACE step was launched is done after the next mtr_server_step;
 * mtr_server_result waits for its hand-over.  Request memory comes from a
 * stream-ordered pool of the server's (plain allocations on a device
 * without memory pools), so a retirement waits for no other device work.
 * A request's result does not depend on what else shares its launches.
 * Times are seconds since mtr_server_create. */
typedef struct mtr_server mtr_server;
int mtr_server_create(int device, int precision, size_t hbm_budget, int max_active, mtr_server **out);
int mtr_server_destroy(mtr_server *srv);
/* Admission order.  MTR_ADMIT_ARRIVAL (the default): oldest request first.
 * MTR_ADMIT_SMALLEST: the pending request with the least work (voxel-view
 * updates of its whole solve, (2SM + 1) N sum_a A_a) first, except that a
 * request that has waited longer than max_wait_s seconds goes first, in
 * arrival order; this shortens the median latency under load and bounds
 * what it costs the largest requests.  The chosen request waits for HBM
 * rather than be overtaken.  Results do not depend on the order. */
enum { MTR_ADMIT_ARRIVAL = 0, MTR_ADMIT_SMALLEST = 1 };
int mtr_server_set_admission(mtr_server *srv, int policy, double max_wait_s);
/* Device bytes a request of this shape holds while active (upper bound). */
int mtr_server_footprint(const mtr_server *srv, const mtr_volume_size *vol, int K,
                         const mtr_projection_size *prj, size_t *bytes);
/* MTR_ERR_OOM when the request alone exceeds the budget; MTR_ERR_NONFINITE
 * for a NaN or Inf in its geometry or projections (checked here, so a
 * queued request always runs). */
int mtr_server_submit(mtr_server *srv, const mtr_volume_size *vol, int K, const mtr_projection_size *prj,
                      const double *const *xc, const double *const *yc, const float *const *p, int M,
                      int S, float lambda, float rho, int64_t *request_id);
/* Completion of the previous step (retirement), admission, and the launch
 * of one MACE step of the active requests; returns while it runs. */
int mtr_server_step(mtr_server *srv, int *completed);
/* done (and step's completed) count finished requests, including failed
 * ones (an admission that could not allocate, for instance).  hbm_in_use
 * still counts a done request until the retirer has freed its memory. */
int mtr_server_status(const mtr_server *srv, int *pending, int *active, int *done, size_t *hbm_in_use);
/* w-bar of a finished request ([Z][Y][X]); times[3] (nullable) receives
 * submission, admission and completion (w-bar on the host) times; waits
 * while the retirer hands the request over.  MTR_ERR_STATE while the
 * request is pending or active; a failed request returns its error. */
int mtr_server_result(mtr_server *srv, int64_t request_id, float *wbar_host, double *times);
int mtr_server_now(const mtr_server *srv, double *seconds);
/* Per-kernel timing of the batched launches (MTR_K_*), as for a context,
 * over the completed steps. */
int mtr_server_timing(mtr_server *srv, int on);
int mtr_server_timing_get(mtr_server *srv, int kernel, int64_t *launches, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* MTR_H_ */
