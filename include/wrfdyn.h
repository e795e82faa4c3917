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
 *  - A handle owns its device workspace and is not thread-safe; one stream
 *    per call.  `stream` is a hipStream_t passed as void* (NULL = default).
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
} wd_state;

/* Create a handle for one grid + namelist; allocates the device workspace on
 * the current HIP device.  Replaces the jit/trace of _rk_scan_step for a
 * static (grid, namelist) pair (operational_mode.py:807-905 cache key). */
int wd_create(const wd_grid *grid, const wd_namelist *nl, wd_handle **out);

/* Advance `n_large_steps` RK3 large steps in place on `stream`
 * (operational_mode.py:3319-3534 per step).  Asynchronous. */
int wd_step(wd_handle *h, const wd_state *state, int n_large_steps, void *stream);

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
