/*
 * wrfdyn: MI355X (gfx950) split-explicit ARW dycore large step, C ABI.
 *
 * This is the drop-in boundary for the reference's dycore step call
 *   _physics_boundary_step(carry, namelist, step_index)   -> OperationalCarry
 *   src/gpuwrf/runtime/operational_mode.py:4996-5013
 * whose core is
 *   _rk_scan_step(carry, namelist)                         -> OperationalCarry
 *   src/gpuwrf/runtime/operational_mode.py:3319-3534
 * with physics and lateral boundaries off (the idealized / dycore-only path).
 * The reference has no FFI for this path (it is Python/JAX); these entry
 * points are what a ctypes / cgo / N-API binding of the step would call.
 *
 * Conventions
 *  - Returns 0 on success or a negative WD_E* code; nothing is thrown across
 *    the ABI.  wd_last_error() returns a message for the last failure.
 *  - All state buffers are caller-owned DEVICE pointers (hipMalloc'd or
 *    torch tensors), FP64, C order (k, j, i), i contiguous -- the reference
 *    layout (src/gpuwrf/contracts/state.py:82-111), so no transposes:
 *      u (nz, ny, nx+1)   v (nz, ny+1, nx)   w, ph_* (nz+1, ny, nx)
 *      theta, p_* (nz, ny, nx)   mu_*, mudf (ny, nx)
 *    They are advanced in place.
 *  - Grid metrics are HOST pointers read once by wd_create.
 *  - A handle owns its device workspace, on the device that was current at
 *    wd_create; calls that launch work need that device current (WD_EINVAL
 *    otherwise).  A handle is not thread-safe and its calls must not overlap
 *    (including a graph or batch replay still running on another stream when
 *    wd_autotune or wd_set_precision drops the graph); one stream per call.
 *    `stream` is a hipStream_t passed as void* (NULL = default).
 */
#ifndef WRFDYN_H
#define WRFDYN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WD_OK 0
#define WD_EINVAL (-1)   /* bad argument / unsupported namelist option */
#define WD_ENOMEM (-2)   /* device allocation failed */
#define WD_EDEVICE (-3)  /* HIP runtime error */

typedef struct wd_handle wd_handle;

/* Grid and metrics: DycoreMetrics (src/gpuwrf/contracts/grid.py:64-245)
 * plus GridSpec.projection dx/dy.  Host pointers. */
typedef struct {
    int nx, ny, nz;
    double dx, dy;
    const double *c1h, *c2h;              /* nz   */
    const double *c1f, *c2f;              /* nz+1 */
    const double *dnw, *rdnw, *dn, *rdn;  /* nz   */
    const double *fnm, *fnp;              /* nz   */
    double cf1, cf2, cf3;
    const double *msftx, *msfty;          /* ny*nx     */
    const double *msfux, *msfuy;          /* ny*(nx+1) */
    const double *msfvx, *msfvy;          /* (ny+1)*nx */
    /* dry-pressure coefficients p = c3 mu + c4 + p_top (DycoreMetrics
     * c3h/c4h/c3f/c4f, contracts/grid.py:84-89), read with hypsometric_opt 2
     * only (NULL otherwise) */
    const double *c3h, *c4h;              /* nz   */
    const double *c3f, *c4f;              /* nz+1 */
    double p_top;
    /* Coriolis parameters on mass points (DycoreMetrics f, e, sina, cosa,
     * contracts/grid.py:65-73): f = 2 Omega sin(lat), e = 2 Omega cos(lat)
     * and the map-projection rotation.  They drive large_step_coriolis
     * (core/rk_addtend_dry.py:364-483) and the cosine-Coriolis term of
     * rw_tend (operational_mode.py:1818-1821).  All four NULL, or f and e
     * zero everywhere, means no rotation (f = e = sina = 0, cosa = 1): the
     * terms vanish and no kernel evaluates them.  f or e given needs all four. */
    const double *f, *e, *sina, *cosa;    /* ny*nx */
} wd_grid;

/* Hot-path fields of OperationalNamelist
 * (src/gpuwrf/runtime/operational_mode.py:412-683). */
typedef struct {
    double dt;              /* dt_s                                   */
    int acoustic_substeps;  /* ns                                     */
    double epssm;
    int top_lid;            /* must be 1                              */
    int w_damping;          /* 0 or 1                                 */
    int damp_opt;           /* 0 or 3                                 */
    double dampcoef;
    double zdamp;
    int diff_6th_opt;       /* 0 off; 1 or 2: the 6th-order horizontal
                               filter, monotonic as the reference applies it */
    double diff_6th_factor; /* diff_6th_factor (reference default 0.12)   */
    int diff_opt;           /* 0 off; 1 with km_opt 4: 2-D Smagorinsky
                               horizontal diffusion (operational_mode.py:2946-2963) */
    int km_opt;             /* 0 or 4                                 */
    double c_s;             /* Smagorinsky constant (reference default 0.25) */
    int scalar_adv_opt;     /* theta advection on the final RK stage: 0 plain
                               h5/v3, 1 positive-definite, 2 monotonic limiter
                               (operational_mode.py:2819-2866) */
    int hypsometric_opt;    /* 1 (or 0) linear al, 2 log-pressure al
                               (rk_addtend_dry.py:196-222, acoustic_wrf.py:394-427) */
    double const_nu_m2_s;   /* > 0: constant-K flux-form diffusion of u, v, w
                               and theta with this K in m^2/s
                               (operational_mode.py:2878-2929); not together
                               with the Smagorinsky option */
    int moist;              /* 1: the state carries the six moisture species
                               (wd_state qv..qg) and they couple into the
                               dynamics: cqu/cqv in the large-step and acoustic
                               u/v pressure gradients, the moist rw_tend and
                               implicit-w cqw (acoustic_wrf.py:303-319,
                               core/advance_w.py:195-290,
                               core/rk_addtend_dry.py:278, :359-360);
                               0: a dry state */
    int moist_adv_opt;      /* with moist: 0 the species are not transported,
                               1 or 2 flux-form h5/v3 transport with the
                               positive-definite (1) or monotonic (2) limiter
                               on the final RK stage and the large-step update
                               (operational_mode.py:3162-3316, :3379-3393) */
} wd_namelist;

/* Prognostic State (src/gpuwrf/contracts/state.py:811-821) plus the carried
 * divergence-damping mass tendency mudf (operational_mode.py:2064,2254).
 * Device pointers, updated in place. */
typedef struct {
    double *u, *v, *w, *theta;
    double *p_total, *p_pert;
    double *ph_total, *ph_pert;
    double *mu_total, *mu_pert;
    double *mudf;
    /* the moisture species (State qv, qc, qr, qi, qs, qg,
     * contracts/state.py:98-117), each nz*ny*nx on mass points: all six with a
     * handle created with moist = 1, all NULL otherwise */
    double *qv, *qc, *qr, *qi, *qs, *qg;
} wd_state;

/* Create a handle for one grid + namelist; allocates the device workspace on
 * the current HIP device.  Replaces the jit/trace of _rk_scan_step for a
 * static (grid, namelist) pair (operational_mode.py:807-905 cache key). */
int wd_create(const wd_grid *grid, const wd_namelist *nl, wd_handle **out);

/* Advance `n_large_steps` RK3 large steps in place on `stream`
 * (operational_mode.py:3319-3534 per step).  Asynchronous. */
int wd_step(wd_handle *h, const wd_state *state, int n_large_steps, void *stream);

/* Graph-captured large step.  The reference compiles its step once per
 * (grid, namelist) and replays the XLA executable (operational_mode.py:807-905
 * jit cache, :3319 _rk_scan_step); the analogue here records the ~80 launches
 * of one large step into a HIP graph and replays that graph.
 * wd_capture_graph records one large step for the buffers in `state`
 * (replacing an earlier capture; profiling must be off).  wd_graph_step
 * replays it n_large_steps times on `stream`, with results identical to
 * wd_step on the same buffers.  The buffers must stay allocated while the
 * graph is in use; wd_autotune drops the captured graph. */
int wd_capture_graph(wd_handle *h, const wd_state *state);
int wd_graph_step(wd_handle *h, int n_large_steps, void *stream);

/* Arithmetic form of the step.  The reference's force_fp64 path is FP64
 * throughout (operational_mode.py:412-683); it is WD_PREC_FP64 here, the
 * default, bit-identical to the CPU oracle.  WD_PREC_MIXED evaluates the
 * flux5/flux3 advection interpolants in FP32 on deviations from the cell's
 * own value, with the mean part (value x transport divergence) added in FP64,
 * and stores the RK tendencies of u, v, w, theta and ph in FP32.  FP64 stays
 * in the large-step pressure gradient, calc_p_rho, advance_w's pressure and
 * buoyancy terms and rw_tend, the tridiagonal solve and the mu/theta
 * accumulation.  Results then differ from the FP64 form by rounding-level
 * amounts; they are not bit-identical.  The mixed form runs the column-march
 * kernels.  Changing the form drops a captured graph. */
#define WD_PREC_FP64 0
#define WD_PREC_MIXED 1
int wd_set_precision(wd_handle *h, int mode);
int wd_get_precision(const wd_handle *h);

/* Tile-shape autotuner (no counterpart in the reference, which leaves launch
 * shapes to XLA).  Times every compiled form of the acoustic column kernel
 * that fits the grid's levels, in XCD-contiguous and in hardware workgroup
 * order, over `trial_steps` large steps each on a scratch copy of `state`
 * (allocated and freed here; the caller's state is not changed), and keeps
 * the fastest for later steps.  Synchronises `stream`.  All forms give
 * identical results.  wd_tuned_config reports the choice: column_tile -1 is
 * the column-march kernel pair, t >= 0 the t-th LDS tile shape.  In the mixed
 * form only the workgroup order is tuned. */
int wd_autotune(wd_handle *h, const wd_state *state, int trial_steps, void *stream);
int wd_tuned_config(const wd_handle *h, int *column_tile, int *xcd_remap);
/* Applies a tuning (a saved wd_tuned_config, or a choice of the caller's):
 * WD_EINVAL for a tile shape that does not hold the grid's levels, or any
 * tile in the mixed form or with moisture, which run the column march.  A
 * captured graph is dropped.  No reference counterpart. */
int wd_set_tuned_config(wd_handle *h, int column_tile, int xcd_remap);

/* Ensembles of independent members with differing grids (no counterpart in
 * the reference, which steps one domain per call).  wd_batch_create captures
 * one large step of every member into one HIP graph: members are spread over
 * `branches` parallel graph branches (largest first, each to the branch with
 * the fewest cells so far) and run back to back within a branch.
 * wd_step_batched launches that graph n_large_steps times on `stream`; each
 * member's result is identical to wd_step on that member alone.  Member
 * handles must be distinct, outlive the batch and not be stepped by other
 * calls while a batched step runs; the state buffers must stay allocated.
 * wd_batch_branch returns the branch a member runs on. */
typedef struct wd_batch wd_batch;
int wd_batch_create(wd_handle *const *members, const wd_state *states, int n_members, int branches,
                    wd_batch **out);
int wd_step_batched(wd_batch *b, int n_large_steps, void *stream);
int wd_batch_branch(const wd_batch *b, int member);
int wd_batch_branches(const wd_batch *b);

/* The same ensemble as one persistent launch per large step (no counterpart
 * in the reference).  One large step of every member is a program of phases
 * (the stage setup's kernels grouped where they touch disjoint arrays, each
 * acoustic kernel, the finish), and a single launch of as many 256-thread
 * workgroups as the GPU holds at once runs it: in every phase the
 * workgroups take the blocks the standalone launches of that phase's
 * kernels would have, for every member, and run the same kernel code on
 * them, then meet at a grid barrier.  The launch is captured in the batch's
 * graph, so wd_step_batched runs it n_large_steps times; each member's
 * result is identical to wd_step on that member alone.  Members must share
 * the precision form and the substep count and use none of the explicit
 * diffusions, the limiter or the moisture (WD_EINVAL).  Each member keeps
 * its tuned workgroup order and, in the FP64 form up to 63 levels, a tuned
 * 16 x 16 LDS column tile; any other column form runs as the column march,
 * which gives the same bits (wd_batch_column_kernel tells which runs).  The grid barrier needs every workgroup resident,
 * so nothing else may run on the GPU during a persistent step: a barrier
 * wait that runs out of polls (wd_batch_set_spin_limit) abandons the step
 * and every later one until wd_batch_status reports it.  wd_batch_branches
 * is 1 and wd_batch_workgroups gives the launch's workgroup count. */
int wd_batch_create_persistent(wd_handle *const *members, const wd_state *states, int n_members,
                               wd_batch **out);
/* Synchronises `stream`; for a persistent batch, WD_EDEVICE if a grid
 * barrier wait ran out since the last call (the steps since then did not
 * complete), after which the flag is cleared and the batch can step again. */
int wd_batch_status(wd_batch *b, void *stream);
/* The bound on one grid barrier wait, in polls of about a microsecond each
 * (default 4,194,304).  Synchronises the device. */
int wd_batch_set_spin_limit(wd_batch *b, unsigned polls);
int wd_batch_workgroups(const wd_batch *b);
/* The acoustic column kernel that `member` runs in the batch's step: -1 for
 * the column march, else the index of its LDS tile shape (as
 * wd_tuned_config).  For a graph batch this is the member's own choice; a
 * persistent batch reports what its step runs.  No reference counterpart: a
 * query of this build's batched forms. */
int wd_batch_column_kernel(const wd_batch *b, int member, int *column_tile);

const char *wd_batch_last_error(const wd_batch *b);
void wd_batch_destroy(wd_batch *b);

/* Adjoint of one large step (the reference has no gradient; this is the
 * vector-Jacobian product of the step wd_step runs, that of
 * operational_mode.py:3319-3534, for the FP64 form).  `start` is the state
 * the step starts from (device buffers, not modified); `bar` holds on entry
 * the adjoint of every array of the state after the step and on return the
 * adjoint of every array of `start`, each of the eleven arrays an
 * independent input.  The step is recomputed from `start` with the forward
 * kernels, and the arrays each adjoint kernel reads are kept on a tape in
 * device memory (wd_adjoint_tape_bytes; allocated on the first call).  Sums of contributions use
 * atomic adds, so results agree with the CPU gradient to rounding, not bit
 * for bit.  The handle's workspace is overwritten.  Asynchronous on
 * `stream`; WD_EINVAL in the mixed form or when a start and a bar buffer
 * coincide.  For n steps, keep the state at the start of every step and
 * call this from the last step back to the first.  It enqueues memsets, so
 * it must not be captured into a graph.
 * Currently switched off: it returns WD_EINVAL with a message, because its
 * first GPU run ended in a memory access fault whose cause is not yet found;
 * the kernel bodies and the whole reverse launch sequence are checked on the
 * host. */
int wd_step_adjoint(wd_handle *h, const wd_state *start, wd_state *bar, void *stream);
size_t wd_adjoint_tape_bytes(const wd_handle *h);

/* A large step that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is
 * wd_create + wd_step + wd_destroy in one call, with the same checks, kernels
 * and results (bit for bit those of wd_step), but every device buffer it
 * needs (the metrics and the workspace) comes from one fixed 64 MiB device
 * array inside the library, one per GPU that calls it, held until the process
 * ends.  That holds the 64x64x40 warm bubble with any one optional term,
 * about 1.8e5 cells; a grid and namelist that need more return WD_EINVAL
 * with the size in the message (use wd_create and wd_step).  No hipMalloc,
 * hipFree or other allocating call is made.  The grid's host arrays are
 * copied in on `stream`, which is synchronised once before the steps are
 * enqueued; the steps themselves run asynchronously.  The array keeps the
 * last call's contents and the workspace is cleared at every call; calls
 * that overlap on one GPU (from several threads or streams) race on it.
 * Errors: wd_last_error(NULL). */
int wd_step_noalloc(const wd_grid *grid, const wd_namelist *nl, const wd_state *state, int n_large_steps,
                    void *stream);

/* Kernel-duration instrumentation for benchmarking (no counterpart in the
 * reference, whose step is one XLA program).  Bit `id` of `kernel_mask`
 * brackets every launch of kernel `id` in subsequent wd_step calls with HIP
 * events on the launch stream; mask 0 disables it.  wd_profile_read waits
 * for the recorded events, writes per-kernel summed milliseconds and launch
 * counts for ids < n, clears the records and returns the number of kernel
 * ids (or a negative WD_E* code). */
int wd_profile_enable(wd_handle *h, unsigned long long kernel_mask);
int wd_profile_read(wd_handle *h, double *total_ms, long long *launches, int n);
int wd_kernel_count(void);
const char *wd_kernel_name(int id);

/* Device workspace bytes the handle holds (for capacity planning). */
size_t wd_workspace_bytes(const wd_handle *h);

/* Message for the last failing call on `h` (or the last wd_create failure
 * when h is NULL).  Never NULL. */
const char *wd_last_error(const wd_handle *h);

void wd_destroy(wd_handle *h);

#ifdef __cplusplus
}
#endif
#endif
