// C ABI of the wrfdyn library (see include/wrfdyn.h).
//
// wd_step runs the reference's large step, _rk_scan_step
// (src/gpuwrf/runtime/operational_mode.py:3319-3534): the RK1 reference is the
// step origin; three stages (dt/3 with one substep at dts=dt/3, dt/2 with
// max(1, ns/2) substeps, dt with ns substeps, both at dts=dt/ns), each
//   stage setup -> acoustic substeps -> finish.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <mutex>
#include <vector>

#include "wrfdyn.h"
#include "wrfdyn_kernels.h"

// Device memory of wd_step_noalloc: the metrics and the workspace of one
// call, at most kScratchBytes.  One copy per GPU that runs the library, held
// until the process ends; in host code `scratch` only names the symbol, and
// its device address comes from hipGetSymbolAddress (scratch_base).
constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

struct wd_handle {
  wd::DGrid g{};
  wd::DNamelist nl{};
  wd::DWork w{};
  void* metric_buf = nullptr;
  void* work_buf = nullptr;
  size_t work_bytes = 0;
  void* diff_buf = nullptr;  // arrays of the optional terms (explicit diffusion, the theta limiter)
  size_t metric_bytes = 0;
  int device = 0;
  int column_tile = 0;
  std::string err;
  wd::Prof prof;
  // wd_capture_graph: one captured large step and the stream it is captured on
  hipStream_t capture_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  // wd_step_adjoint: the tape of one large step (allocated on first use)
  wd::AdjTape* tape = nullptr;
  size_t tape_bytes = 0;
};

// An ensemble: member handles (each its own grid, namelist and workspace)
// whose large steps are captured together into one graph, spread over
// `branches` parallel branches (one capture stream each).
struct wd_batch {
  std::vector<wd_handle*> members;
  std::vector<wd::DState> states;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<int> branch_of;  // member -> branch
  int branches = 1;            // branches in use (<= streams.size())
  hipGraphExec_t graph = nullptr;
  int device = 0;
  std::string err;
  // wd_batch_create_persistent: the graph holds one launch of the persistent
  // step over `grid` workgroups; its member table, program and grid barrier
  // share one device allocation
  bool persistent = false;
  unsigned grid = 0;
  std::vector<int> column_kernel;  // per member: kColumnMarch or the tile it runs
  void* pbuf = nullptr;
  wd::PBarrier* bar = nullptr;
};

namespace {

thread_local std::string g_create_error;

int fail(wd_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(wd_handle* h, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(h, WD_EDEVICE, m);
}

// Launches go to the current device; a handle's workspace lives on the device
// it was created on.
int check_device(wd_handle* h) {
  int cur = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(h, e, "hipGetDevice");
  if (cur != h->device)
    return fail(h, WD_EINVAL, "the current HIP device (" + std::to_string(cur) + ") is not the handle's device (" +
                                  std::to_string(h->device) + ")");
  return WD_OK;
}

// moist: whether the state must carry the six species (-1: the handle's namelist says)
int check_state(wd_handle* h, const wd_state* st, wd::DState* out, int moist_flag = -1) {
  if (!st) return fail(h, WD_EINVAL, "state is NULL");
  double* ptrs[] = {st->u, st->v, st->w, st->theta, st->p_total, st->p_pert, st->ph_total,
                    st->ph_pert, st->mu_total, st->mu_pert, st->mudf};
  for (double* p : ptrs)
    if (!p) return fail(h, WD_EINVAL, "a state pointer is NULL");
  *out = wd::DState{st->u, st->v, st->w, st->theta, st->p_total, st->p_pert,
                    st->ph_total, st->ph_pert, st->mu_total, st->mu_pert, st->mudf};
  // the moisture species: all six for a moist handle, none for a dry one
  double* q[wd::kNumMoist] = {st->qv, st->qc, st->qr, st->qi, st->qs, st->qg};
  int nq = 0;
  for (double* p : q) nq += p != nullptr;
  const bool moist = moist_flag >= 0 ? moist_flag != 0 : (h && h->nl.moist);
  if (moist && nq != wd::kNumMoist)
    return fail(h, WD_EINVAL, "a moist handle needs all six moisture species (qv, qc, qr, qi, qs, qg)");
  if (!moist && nq != 0)
    return fail(h, WD_EINVAL, "the state carries moisture species but the handle was created with moist = 0");
  if (moist)
    for (int sp = 0; sp < wd::kNumMoist; ++sp) out->q[sp] = q[sp];
  return WD_OK;
}

// Enqueue n large steps of _rk_scan_step (operational_mode.py:3319-3534) on
// stream s: no host synchronisation, allocation or memset, so the sequence
// can also be captured into a graph (profiling records events, so it is
// refused during capture).
int enqueue_steps(wd_handle* h, const wd::DState& ds, int n_large_steps, hipStream_t s) {
  const wd::DGrid& g = h->g;
  const wd::DWork& w = h->w;
  const size_t M = (size_t)g.nz * g.ny * g.nx, Wn = (size_t)(g.nz + 1) * g.ny * g.nx;
  const size_t U = (size_t)g.nz * g.ny * (g.nx + 1), V = (size_t)g.nz * (g.ny + 1) * g.nx;
  const size_t C = (size_t)g.ny * g.nx;
  const double dt = h->nl.dt;
  const int ns = h->nl.ns;
  struct StageDesc {
    int rk_step;
    double dts;
    int nsub;
  };
  const StageDesc stages[3] = {
      {1, dt / 3.0, 1},
      {2, dt / (double)ns, ns / 2 > 1 ? ns / 2 : 1},
      {3, dt / (double)ns, ns},
  };
  for (int step = 0; step < n_large_steps; ++step) {
    hipError_t e;
    wd::Prof* prof = h->prof.mask ? &h->prof : nullptr;
    const std::pair<double*, const double*> copies[] = {
        {w.u0, ds.u}, {w.v0, ds.v}, {w.w0, ds.w}, {w.th0, ds.theta}, {w.ph0, ds.ph_pert}, {w.mu0, ds.mu_pert}};
    const size_t sizes[] = {U, V, Wn, M, Wn, C};
    size_t tok = prof ? prof->begin(wd::KID_ORIGIN_COPY, s) : 0;
    for (int q = 0; q < 6; ++q) {
      e = hipMemcpyAsync(copies[q].first, copies[q].second, sizes[q] * sizeof(double), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return hip_fail(h, e, "hipMemcpyAsync(origin)");
    }
    if (w.mut0) {
      e = hipMemcpyAsync(w.mut0, ds.mu_total, C * sizeof(double), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return hip_fail(h, e, "hipMemcpyAsync(origin)");
    }
    if (w.q0[0])  // the species at the step origin (transport on)
      for (int sp = 0; sp < wd::kNumMoist; ++sp) {
        e = hipMemcpyAsync(w.q0[sp], ds.q[sp], M * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(h, e, "hipMemcpyAsync(origin)");
      }
    if (prof) prof->end(tok, s);
    for (const StageDesc& sd : stages) {
      e = wd::launch_stage_setup(g, h->nl, ds, w, sd.rk_step, sd.dts, s, prof);
      if (e != hipSuccess) return hip_fail(h, e, "stage setup launch");
      for (int it = 0; it < sd.nsub; ++it) {
        e = wd::launch_acoustic_substep(g, h->nl, ds, w, sd.dts, it == 0, h->column_tile, s, prof);
        if (e != hipSuccess) return hip_fail(h, e, "acoustic substep launch");
      }
      e = wd::launch_stage_finish(g, ds, w, s, prof);
      if (e != hipSuccess) return hip_fail(h, e, "stage finish launch");
      // the species' update with the stage's dt_rk: dt/3, dt/2, dt
      // (operational_mode.py:3527-3529)
      if (w.q0[0]) {
        const double dt_rk = sd.rk_step == 1 ? dt / 3.0 : (sd.rk_step == 2 ? 0.5 * dt : dt);
        e = wd::launch_moist_update(g, ds, w, dt_rk, s, prof);
        if (e != hipSuccess) return hip_fail(h, e, "moisture update launch");
      }
    }
  }
  return WD_OK;
}

// Abandon an open capture on s (after a failed enqueue) without leaking the graph.
void abort_capture(hipStream_t s) {
  hipGraph_t partial = nullptr;
  if (hipStreamEndCapture(s, &partial) == hipSuccess && partial) hipGraphDestroy(partial);
}

size_t state_doubles(const wd::DGrid& g, int field) {
  const size_t nx = g.nx, ny = g.ny, nz = g.nz;
  switch (field) {
    case 0: return nz * ny * (nx + 1);                  // u
    case 1: return nz * (ny + 1) * nx;                  // v
    case 2: case 6: case 7: return (nz + 1) * ny * nx;  // w, ph_total, ph_pert
    case 3: case 4: case 5: return nz * ny * nx;        // theta, p_total, p_pert
    case 8: case 9: case 10: return ny * nx;            // mu_total, mu_pert, mudf
    default: return nz * ny * nx;                       // the moisture species
  }
}

// fields 0..10 of the dry state, then the six moisture species
double** state_field(wd::DState& s, int field) {
  double** f[] = {&s.u, &s.v, &s.w, &s.theta, &s.p_total, &s.p_pert,
                  &s.ph_total, &s.ph_pert, &s.mu_total, &s.mu_pert, &s.mudf,
                  &s.q[0], &s.q[1], &s.q[2], &s.q[3], &s.q[4], &s.q[5]};
  return f[field];
}
constexpr int kStateFields = 11;
int state_fields(const wd::DNamelist& nl) { return nl.moist ? kStateFields + wd::kNumMoist : kStateFields; }

}  // namespace

namespace {

// Longest-processing-time assignment of members to nb branches: members by
// decreasing cell count, each to the branch with the fewest cells so far.
std::vector<int> assign_branches(const std::vector<wd_handle*>& members, int nb) {
  const int n = (int)members.size();
  std::vector<int> order(n), branch_of(n, 0);
  for (int m = 0; m < n; ++m) order[m] = m;
  auto cells = [&](int m) {
    const wd::DGrid& g = members[m]->g;
    return (double)g.nx * g.ny * g.nz;
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cells(a) > cells(c); });
  std::vector<double> load(nb, 0.0);
  for (int m : order) {
    int best = 0;
    for (int k = 1; k < nb; ++k)
      if (load[k] < load[best]) best = k;
    branch_of[m] = best;
    load[best] += cells(m);
  }
  return branch_of;
}

// Capture one large step of every member on `states` into *out: branch 0
// forks to the others, each branch runs its members back to back, and
// branch 0 joins them.  Errors are reported through g_create_error.
int capture_batch(wd_batch* b, const std::vector<wd::DState>& states, const std::vector<int>& branch_of, int nb,
                  hipGraphExec_t* out) {
  hipStream_t s0 = b->streams[0];
  hipError_t r = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (r != hipSuccess) return hip_fail(nullptr, r, "hipStreamBeginCapture");
  r = hipEventRecord(b->events[0], s0);
  for (int k = 1; r == hipSuccess && k < nb; ++k) r = hipStreamWaitEvent(b->streams[k], b->events[0], 0);
  if (r != hipSuccess) {
    abort_capture(s0);
    return hip_fail(nullptr, r, "capture fork");
  }
  const int n = (int)b->members.size();
  for (int k = 0; k < nb; ++k)
    for (int m = 0; m < n; ++m) {
      if (branch_of[m] != k) continue;
      int rc = enqueue_steps(b->members[m], states[m], 1, b->streams[k]);
      if (rc != WD_OK) {
        abort_capture(s0);
        return fail(nullptr, rc, "member " + std::to_string(m) + ": " + b->members[m]->err);
      }
    }
  for (int k = 1; r == hipSuccess && k < nb; ++k) {
    r = hipEventRecord(b->events[k], b->streams[k]);
    if (r == hipSuccess) r = hipStreamWaitEvent(s0, b->events[k], 0);
  }
  if (r != hipSuccess) {
    abort_capture(s0);
    return hip_fail(nullptr, r, "capture join");
  }
  hipGraph_t graph = nullptr;
  r = hipStreamEndCapture(s0, &graph);
  if (r != hipSuccess) return hip_fail(nullptr, r, "hipStreamEndCapture");
  r = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (r != hipSuccess) return hip_fail(nullptr, r, "hipGraphInstantiate");
  return WD_OK;
}


int validate_create(const wd_grid* grid, const wd_namelist* nl, bool* rot_out) {
  if (!grid || !nl) return fail(nullptr, WD_EINVAL, "grid or namelist is NULL");
  const int nx = grid->nx, ny = grid->ny, nz = grid->nz;
  if (nx < 4 || ny < 1 || nz < 4)
    return fail(nullptr, WD_EINVAL, "grid too small: need nx >= 4, ny >= 1, nz >= 4");
  {
    // the k-marching kernels address each 3-D array with 32-bit byte offsets
    const double cells_w = (double)(nz + 1) * (double)(ny + 1) * (double)(nx + 1);
    if (cells_w * 8.0 > 4294967295.0)
      return fail(nullptr, WD_EINVAL, "grid too large: each 3-D field must stay below 4 GiB");
  }
  if (!(grid->dx > 0.0) || !(grid->dy > 0.0)) return fail(nullptr, WD_EINVAL, "dx and dy must be positive");
  const double* req[] = {grid->c1h, grid->c2h, grid->c1f, grid->c2f, grid->dnw, grid->rdnw, grid->dn,
                         grid->rdn, grid->fnm, grid->fnp, grid->msftx, grid->msfty, grid->msfux,
                         grid->msfuy, grid->msfvx, grid->msfvy};
  for (const double* p : req)
    if (!p) return fail(nullptr, WD_EINVAL, "a grid metric pointer is NULL");
  if (nl->acoustic_substeps < 1) return fail(nullptr, WD_EINVAL, "acoustic_substeps must be >= 1");
  if (!(nl->dt > 0.0)) return fail(nullptr, WD_EINVAL, "dt must be positive");
  if (nl->top_lid != 1) return fail(nullptr, WD_EINVAL, "only top_lid=1 (rigid lid) is implemented");
  if (nl->w_damping != 0 && nl->w_damping != 1) return fail(nullptr, WD_EINVAL, "w_damping must be 0 or 1");
  if (nl->damp_opt != 0 && nl->damp_opt != 3) return fail(nullptr, WD_EINVAL, "damp_opt must be 0 or 3");
  if (nl->damp_opt == 3 && nl->dampcoef > 0.0 && !(nl->zdamp > 0.0))
    return fail(nullptr, WD_EINVAL, "zdamp must be positive with damp_opt=3");
  if (nl->diff_6th_opt < 0 || nl->diff_6th_opt > 2) return fail(nullptr, WD_EINVAL, "diff_6th_opt must be 0, 1 or 2");
  if ((nl->diff_opt != 0 && nl->diff_opt != 1) || (nl->km_opt != 0 && nl->km_opt != 4) ||
      (nl->diff_opt == 1 && nl->km_opt != 4))
    return fail(nullptr, WD_EINVAL,
                "diffusion: only diff_opt=0, or diff_opt=1 with km_opt=4 (2-D Smagorinsky), is implemented");
  if (!(nl->c_s >= 0.0)) return fail(nullptr, WD_EINVAL, "c_s must be >= 0");
  if (!(nl->const_nu_m2_s >= 0.0 && nl->const_nu_m2_s <= 1.0e300))
    return fail(nullptr, WD_EINVAL, "const_nu_m2_s must be finite and >= 0");
  if (nl->const_nu_m2_s > 0.0 && nl->diff_opt == 1)
    return fail(nullptr, WD_EINVAL, "const_nu_m2_s > 0 and the 2-D Smagorinsky diffusion do not run together");
  if (nl->hypsometric_opt < 0 || nl->hypsometric_opt > 2)
    return fail(nullptr, WD_EINVAL, "hypsometric_opt must be 1 (linear) or 2 (log-pressure)");
  if (nl->hypsometric_opt == 2 && (!grid->c3h || !grid->c4h || !grid->c3f || !grid->c4f))
    return fail(nullptr, WD_EINVAL, "hypsometric_opt=2 needs the c3h, c4h, c3f and c4f metrics");
  if (nl->scalar_adv_opt < 0 || nl->scalar_adv_opt > 2)
    return fail(nullptr, WD_EINVAL, "scalar_adv_opt must be 0, 1 (positive-definite) or 2 (monotonic)");
  if (nl->moist != 0 && nl->moist != 1) return fail(nullptr, WD_EINVAL, "moist must be 0 (dry) or 1");
  if (nl->moist_adv_opt < 0 || nl->moist_adv_opt > 2)
    return fail(nullptr, WD_EINVAL, "moist_adv_opt must be 0 (off), 1 (positive-definite) or 2 (monotonic)");
  if (nl->moist_adv_opt != 0 && !nl->moist)
    return fail(nullptr, WD_EINVAL, "moist_adv_opt transports the moisture: it needs moist = 1");
  // rotation: all four Coriolis arrays or none; it runs when f or e is nonzero
  bool& rot = *rot_out;
  rot = false;
  if (grid->f || grid->e || grid->sina || grid->cosa) {
    if (!grid->f || !grid->e || !grid->sina || !grid->cosa)
      return fail(nullptr, WD_EINVAL, "rotation needs all four of f, e, sina and cosa (or none)");
    for (size_t q = 0; q < (size_t)ny * nx && !rot; ++q) rot = grid->f[q] != 0.0 || grid->e[q] != 0.0;
  }

  return WD_OK;
}

// The scalar parameters of the grid and namelist (everything but the device
// pointers of the metrics).
void init_params(const wd_grid* grid, const wd_namelist* nl, bool rot, wd::DGrid& g, wd::DNamelist& n) {
  const int nz = grid->nz;
  g.nx = grid->nx;
  g.ny = grid->ny;
  g.nz = nz;
  g.dx = grid->dx;
  g.dy = grid->dy;
  g.rdx = 1.0 / grid->dx;
  g.rdy = 1.0 / grid->dy;
  g.rot = rot ? 1 : 0;
  g.p_top = grid->p_top;
  g.hypsometric_opt = nl->hypsometric_opt;
  g.cf1 = grid->cf1;
  g.cf2 = grid->cf2;
  g.cf3 = grid->cf3;
  // WRF top-face extrapolation weights cfn/cfn1 (core/rk_addtend_dry.py:305-312).
  double dn_top = grid->dn[nz - 1];
  double dn_safe = std::fabs(dn_top) > 1.0e-30 ? dn_top : 1.0;
  g.cfn = (0.5 * grid->dnw[nz - 1] + grid->dn[nz - 1]) / dn_safe;
  g.cfn1 = -0.5 * grid->dnw[nz - 1] / dn_safe;
  // XCD-contiguous tile order (default on); WRFDYN_XCD_REMAP=0 keeps the
  // hardware dispatch order, for measurement.
  g.xcd_remap = 1;
  if (const char* env = std::getenv("WRFDYN_XCD_REMAP")) g.xcd_remap = std::atoi(env) != 0;

  n.dt = nl->dt;
  n.ns = nl->acoustic_substeps;
  n.epssm = nl->epssm;
  n.top_lid = nl->top_lid;
  n.w_damping = nl->w_damping;
  n.damp_opt = nl->damp_opt;
  n.dampcoef = nl->dampcoef;
  n.zdamp = nl->zdamp;
  n.diff_6th_opt = nl->diff_6th_opt;
  n.diff_6th_factor = nl->diff_6th_factor;
  n.smag = nl->diff_opt == 1 && nl->km_opt == 4;
  n.c_s = nl->c_s;
  n.scalar_adv_opt = nl->scalar_adv_opt;
  n.const_nu = nl->const_nu_m2_s;
  n.moist = nl->moist;
  n.moist_adv_opt = nl->moist ? nl->moist_adv_opt : 0;
  n.mixed = 0;
}

struct WsSlot {
  double** p;
  size_t n;  // doubles
};
constexpr int kMaxWsSlots = 64;
size_t slot_bytes(size_t n) { return (n * sizeof(double) + 255) / 256 * 256; }

// The workspace arrays of a handle: the base set (optional = false), or the
// arrays of the optional terms the namelist turns on (optional = true): the
// filter's tendencies, the Smagorinsky viscosity and tendencies or the
// constant-K tendencies and layer sums, the limiter's origin mass, scales
// and face fluxes.  Returns the count written to out.
int workspace_slots(const wd::DGrid& g, const wd::DNamelist& n, bool optional, wd::DWork& w, WsSlot* out) {
  const size_t nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t M = nz * ny * nx, Wn = (nz + 1) * ny * nx, U = nz * ny * (nx + 1), V = nz * (ny + 1) * nx;
  const size_t c = ny * nx, cu = ny * (nx + 1), cv = (ny + 1) * nx;
  int k = 0;
  auto add = [&](double** p, size_t len) { out[k++] = WsSlot{p, len}; };
  if (!optional) {
    const WsSlot base[] = {
        {&w.u0, U},       {&w.v0, V},        {&w.w0, Wn},      {&w.th0, M},      {&w.ph0, Wn},
        {&w.mu0, c},      {&w.ru, M},        {&w.rv, M},       {&w.rom, Wn},     {&w.ut, U},
        {&w.vt, V},       {&w.wt, Wn},       {&w.tht, M},      {&w.alt_raw, M},  {&w.al_ls, M},
        {&w.php_ls, M},   {&w.mub, c},       {&w.mut, c},      {&w.muts0, c},    {&w.mu_work0, c},
        {&w.mu, c},       {&w.muts, c},      {&w.muave, c},    {&w.muu, cu},     {&w.muus, cu},
        {&w.muv, cv},     {&w.muvs, cv},     {&w.u_work, U},   {&w.v_work, V},   {&w.w_work, Wn},
        {&w.th_work, M},  {&w.ph_work, Wn},  {&w.c2a, M},      {&w.alt, M},      {&w.pb, M},
        {&w.ph1t, Wn},    {&w.php, M},       {&w.p, M},        {&w.al, M},       {&w.pm1, M},
        {&w.rw_tend, Wn}, {&w.ph_tend, Wn},  {&w.ca, Wn},      {&w.calpha, Wn},  {&w.cgamma, Wn},
        {&w.ww, Wn},      {&w.t2ave, M},     {&w.rhs, Wn},     {&w.wfwd, Wn},
    };
    for (const WsSlot& s : base) add(s.p, s.n);
    return k;
  }
  if (n.diff_6th_opt != 0) {
    add(&w.d6u, U);
    add(&w.d6v, V);
    add(&w.d6w, Wn);
    add(&w.d6t, M);
  }
  if (n.smag) {
    add(&w.xkmh, M);
    add(&w.smu, U);
    add(&w.smv, V);
    add(&w.smw, Wn);
    add(&w.smt, M);
  }
  if (n.const_nu > 0.0) {
    add(&w.smu, U);
    add(&w.smv, V);
    add(&w.smw, Wn);
    add(&w.smt, M);
    add(&w.ck_col, c);
    add(&w.ck_dz, 1);
  }
  if (n.moist) {
    add(&w.cqu, U);
    add(&w.cqv, V);
    add(&w.cqwr, Wn);
    if (n.moist_adv_opt != 0)
      for (int sp = 0; sp < wd::kNumMoist; ++sp) {
        add(&w.q0[sp], M);
        add(&w.qt[sp], M);
      }
  }
  // the limiter's arrays serve theta's and the species' limited advection
  if (n.scalar_adv_opt != 0 || n.moist_adv_opt != 0) {
    add(&w.mut0, c);
    add(&w.lim_sa, M);
    add(&w.lim_sb, M);
    add(&w.lim_ax, M);
    add(&w.lim_lx, M);
    add(&w.lim_ay, M);
    add(&w.lim_ly, M);
    add(&w.lim_az, Wn);
    add(&w.lim_lz, Wn);
  }
  return k;
}

// The device address of this GPU's `scratch`, looked up once per device with
// that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_base(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, WD_EINVAL, "HIP device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return WD_OK;
}

__global__ void __launch_bounds__(256) k_fill(double* p, size_t n, double v) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = v;
}

}  // namespace

extern "C" {

int wd_create(const wd_grid* grid, const wd_namelist* nl, wd_handle** out) {
  if (!out) return fail(nullptr, WD_EINVAL, "out is NULL");
  *out = nullptr;
  bool rot = false;
  const int vrc = validate_create(grid, nl, &rot);
  if (vrc != WD_OK) return vrc;
  const int nx = grid->nx, ny = grid->ny, nz = grid->nz;

  wd_handle* h = new (std::nothrow) wd_handle();
  if (!h) return fail(nullptr, WD_ENOMEM, "host allocation failed");
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(nullptr, e, "hipGetDevice");
  }

  // ---- metrics: 1-D vertical + 2-D map factors in one device buffer ----
  const size_t n1 = (size_t)nz, n1f = (size_t)nz + 1;
  const size_t c = (size_t)ny * nx, cu = (size_t)ny * (nx + 1), cv = (size_t)(ny + 1) * nx;
  std::vector<double> host;
  host.reserve(10 * n1 + 4 * n1f + 2 * c + 2 * cu + 2 * cv);
  std::vector<size_t> offs;
  auto put = [&](const double* src, size_t n) {
    offs.push_back(host.size());
    host.insert(host.end(), src, src + n);
  };
  put(grid->c1h, n1);
  put(grid->c2h, n1);
  put(grid->c1f, n1f);
  put(grid->c2f, n1f);
  put(grid->dnw, n1);
  put(grid->rdnw, n1);
  put(grid->dn, n1);
  put(grid->rdn, n1);
  put(grid->fnm, n1);
  put(grid->fnp, n1);
  put(grid->msftx, c);
  put(grid->msfty, c);
  put(grid->msfux, cu);
  put(grid->msfuy, cu);
  put(grid->msfvx, cv);
  put(grid->msfvy, cv);
  // the dry-pressure coefficients; without them (hypsometric_opt 1) the
  // idealized c3 = 1, c4 = 0, which nothing reads
  const std::vector<double> one_h(n1, 1.0), zero_h(n1, 0.0), one_f(n1f, 1.0), zero_f(n1f, 0.0);
  put(grid->c3h ? grid->c3h : one_h.data(), n1);
  put(grid->c4h ? grid->c4h : zero_h.data(), n1);
  put(grid->c3f ? grid->c3f : one_f.data(), n1f);
  put(grid->c4f ? grid->c4f : zero_f.data(), n1f);
  if (rot) {
    put(grid->f, c);
    put(grid->e, c);
    put(grid->sina, c);
    put(grid->cosa, c);
  }
  h->metric_bytes = host.size() * sizeof(double);
  e = hipMalloc(&h->metric_buf, h->metric_bytes);
  if (e != hipSuccess) {
    delete h;
    return fail(nullptr, WD_ENOMEM, std::string("hipMalloc(metrics): ") + hipGetErrorString(e));
  }
  e = hipMemcpy(h->metric_buf, host.data(), h->metric_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(h->metric_buf);
    delete h;
    return hip_fail(nullptr, e, "hipMemcpy(metrics)");
  }
  const double* mb = static_cast<const double*>(h->metric_buf);
  wd::DGrid& g = h->g;
  int o = 0;
  g.c1h = mb + offs[o++];
  g.c2h = mb + offs[o++];
  g.c1f = mb + offs[o++];
  g.c2f = mb + offs[o++];
  g.dnw = mb + offs[o++];
  g.rdnw = mb + offs[o++];
  g.dn = mb + offs[o++];
  g.rdn = mb + offs[o++];
  g.fnm = mb + offs[o++];
  g.fnp = mb + offs[o++];
  g.msftx = mb + offs[o++];
  g.msfty = mb + offs[o++];
  g.msfux = mb + offs[o++];
  g.msfuy = mb + offs[o++];
  g.msfvx = mb + offs[o++];
  g.msfvy = mb + offs[o++];
  g.c3h = mb + offs[o++];
  g.c4h = mb + offs[o++];
  g.c3f = mb + offs[o++];
  g.c4f = mb + offs[o++];
  g.fcor = rot ? mb + offs[o++] : nullptr;
  g.ecor = rot ? mb + offs[o++] : nullptr;
  g.sina = rot ? mb + offs[o++] : nullptr;
  g.cosa = rot ? mb + offs[o++] : nullptr;
  init_params(grid, nl, rot, g, h->nl);

  // ---- workspace ----
  wd::DWork& w = h->w;
  WsSlot slots[kMaxWsSlots];
  const int nslots = workspace_slots(g, h->nl, false, w, slots);
  // 256-byte aligned sub-buffers
  size_t total = 0;
  for (int q = 0; q < nslots; ++q) total += slot_bytes(slots[q].n);
  e = hipMalloc(&h->work_buf, total);
  if (e != hipSuccess) {
    hipFree(h->metric_buf);
    delete h;
    return fail(nullptr, WD_ENOMEM, std::string("hipMalloc(workspace): ") + hipGetErrorString(e));
  }
  h->work_bytes = total;
  char* base = static_cast<char*>(h->work_buf);
  size_t off = 0;
  for (int q = 0; q < nslots; ++q) {
    *slots[q].p = reinterpret_cast<double*>(base + off);
    off += slot_bytes(slots[q].n);
  }
  e = hipMemset(h->work_buf, 0, total);
  if (e != hipSuccess) {
    hipFree(h->work_buf);
    hipFree(h->metric_buf);
    delete h;
    return hip_fail(nullptr, e, "hipMemset(workspace)");
  }
  h->nl.mixed = 0;
  WsSlot dslots[kMaxWsSlots];
  const int ndslots = workspace_slots(g, h->nl, true, w, dslots);
  if (ndslots > 0) {
    // the arrays of the optional terms, each 256-byte aligned: the filter's
    // tendencies, the Smagorinsky viscosity and tendencies, the limiter's
    // origin mass and scales
    size_t d_total = 0;
    for (int q = 0; q < ndslots; ++q) d_total += slot_bytes(dslots[q].n);
    e = hipMalloc(&h->diff_buf, d_total);
    if (e != hipSuccess) {
      hipFree(h->work_buf);
      hipFree(h->metric_buf);
      delete h;
      return fail(nullptr, WD_ENOMEM, std::string("hipMalloc(optional terms): ") + hipGetErrorString(e));
    }
    size_t od = 0;
    for (int q = 0; q < ndslots; ++q) {
      *dslots[q].p = reinterpret_cast<double*>(static_cast<char*>(h->diff_buf) + od);
      od += slot_bytes(dslots[q].n);
    }
    h->work_bytes += d_total;
  }
  // Acoustic column kernel: the column-march pair (default) or, with
  // WRFDYN_COLUMN_TILE="TXxKT", the LDS-tiled kernel in one of its compiled
  // tile shapes (when that shape holds nz + 1 levels).
  h->column_tile = wd::kColumnMarch;
  if (const char* env = std::getenv("WRFDYN_COLUMN_TILE")) {
    int tx = 0, kt = 0;
    if (std::strcmp(env, "march") == 0) h->column_tile = wd::kColumnMarch;
    else if (std::sscanf(env, "%dx%d", &tx, &kt) == 2) {
      for (int t = 0; t < wd::kNumColumnTiles; ++t)
        if (wd::kColumnTiles[t].tx == tx && wd::kColumnTiles[t].kt == kt &&
            nz + 1 <= kt * wd::kMaxLevelsPerThread)
          h->column_tile = t;
    }
  }
  // the LDS-tiled kernel has no moisture factor
  if (h->nl.moist) h->column_tile = wd::kColumnMarch;
  *out = h;
  return WD_OK;
}

int wd_step(wd_handle* h, const wd_state* st, int n_large_steps, void* stream) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (n_large_steps < 0) return fail(h, WD_EINVAL, "n_large_steps must be >= 0");
  wd::DState ds;
  int rc = check_state(h, st, &ds);
  if (rc == WD_OK) rc = check_device(h);
  if (rc != WD_OK) return rc;
  return enqueue_steps(h, ds, n_large_steps, static_cast<hipStream_t>(stream));
}

int wd_capture_graph(wd_handle* h, const wd_state* st) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (h->prof.mask) return fail(h, WD_EINVAL, "disable profiling (wd_profile_enable(h, 0)) before capturing");
  wd::DState ds;
  int rc = check_state(h, st, &ds);
  if (rc == WD_OK) rc = check_device(h);
  if (rc != WD_OK) return rc;
  hipError_t e;
  if (!h->capture_stream) {
    e = hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(h, e, "hipStreamCreateWithFlags");
  }
  e = hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(h, e, "hipStreamBeginCapture");
  rc = enqueue_steps(h, ds, 1, h->capture_stream);
  if (rc != WD_OK) {
    abort_capture(h->capture_stream);
    return rc;
  }
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(h->capture_stream, &graph);
  if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
  if (h->graph) hipGraphExecDestroy(h->graph);
  h->graph = exec;
  return WD_OK;
}

int wd_graph_step(wd_handle* h, int n_large_steps, void* stream) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (!h->graph) return fail(h, WD_EINVAL, "no captured graph: call wd_capture_graph first");
  if (n_large_steps < 0) return fail(h, WD_EINVAL, "n_large_steps must be >= 0");
  if (int rc = check_device(h); rc != WD_OK) return rc;
  for (int i = 0; i < n_large_steps; ++i) {
    hipError_t e = hipGraphLaunch(h->graph, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(h, e, "hipGraphLaunch");
  }
  return WD_OK;
}

int wd_autotune(wd_handle* h, const wd_state* st, int trial_steps, void* stream) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (trial_steps < 1) return fail(h, WD_EINVAL, "trial_steps must be >= 1");
  if (h->prof.mask) return fail(h, WD_EINVAL, "disable profiling before autotuning");
  wd::DState src;
  int rc = check_state(h, st, &src);
  if (rc == WD_OK) rc = check_device(h);
  if (rc != WD_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // scratch copy of the state: trials advance it, the caller's state is untouched
  size_t total = 0;
  const int nfields = state_fields(h->nl);
  for (int f = 0; f < nfields; ++f) total += (state_doubles(h->g, f) * sizeof(double) + 255) / 256 * 256;
  void* buf = nullptr;
  hipError_t e = hipMalloc(&buf, total);
  if (e != hipSuccess) return fail(h, WD_ENOMEM, std::string("hipMalloc(autotune state): ") + hipGetErrorString(e));
  wd::DState scratch = src;
  {
    size_t off = 0;
    for (int f = 0; f < nfields; ++f) {
      *state_field(scratch, f) = reinterpret_cast<double*>(static_cast<char*>(buf) + off);
      off += (state_doubles(h->g, f) * sizeof(double) + 255) / 256 * 256;
    }
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  const int keep_tile = h->column_tile, keep_remap = h->g.xcd_remap;
  // candidates: the column-march pair and every LDS tile shape that holds nz + 1 levels, each in
  // both workgroup orders
  std::vector<std::pair<int, int>> cand;
  for (int remap = 1; remap >= 0; --remap) {
    cand.emplace_back(wd::kColumnMarch, remap);
    for (int t = 0; t < wd::kNumColumnTiles && !h->nl.mixed && !h->nl.moist; ++t)
      if (h->g.nz + 1 <= wd::kColumnTiles[t].kt * wd::kMaxLevelsPerThread) cand.emplace_back(t, remap);
  }
  int best = -1;
  float best_ms = 0.0f;
  auto reset = [&]() -> hipError_t {
    for (int f = 0; f < nfields; ++f) {
      hipError_t r = hipMemcpyAsync(*state_field(scratch, f), *state_field(src, f),
                                    state_doubles(h->g, f) * sizeof(double), hipMemcpyDeviceToDevice, s);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  };
  e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  for (size_t c = 0; e == hipSuccess && rc == WD_OK && c < cand.size(); ++c) {
    h->column_tile = cand[c].first;
    h->g.xcd_remap = cand[c].second;
    e = reset();
    if (e != hipSuccess) break;
    rc = enqueue_steps(h, scratch, 1, s);  // warm-up (first touch, code load)
    if (rc != WD_OK) break;
    e = reset();
    if (e == hipSuccess) e = hipEventRecord(ev[0], s);
    if (e != hipSuccess) break;
    rc = enqueue_steps(h, scratch, trial_steps, s);
    if (rc != WD_OK) break;
    e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (e == hipSuccess && (best < 0 || ms < best_ms)) {
      best = (int)c;
      best_ms = ms;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  for (hipEvent_t x : ev)
    if (x) hipEventDestroy(x);
  hipFree(buf);
  if (e != hipSuccess || rc != WD_OK || best < 0) {
    h->column_tile = keep_tile;
    h->g.xcd_remap = keep_remap;
    if (rc != WD_OK) return rc;
    return hip_fail(h, e, "autotune trial");
  }
  h->column_tile = cand[best].first;
  h->g.xcd_remap = cand[best].second;
  if (h->graph) {  // a captured graph holds the old launch shapes
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return WD_OK;
}

int wd_set_tuned_config(wd_handle* h, int column_tile, int xcd_remap) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (xcd_remap != 0 && xcd_remap != 1) return fail(h, WD_EINVAL, "xcd_remap must be 0 or 1");
  if (column_tile != wd::kColumnMarch) {
    if (column_tile < 0 || column_tile >= wd::kNumColumnTiles)
      return fail(h, WD_EINVAL, "column_tile must be -1 or a compiled tile shape index");
    if (h->g.nz + 1 > wd::kColumnTiles[column_tile].kt * wd::kMaxLevelsPerThread)
      return fail(h, WD_EINVAL, "the tile shape does not hold the grid's levels");
    if (h->nl.mixed || h->nl.moist)
      return fail(h, WD_EINVAL, "the mixed form and moist states run the column march (column_tile -1)");
  }
  h->column_tile = column_tile;
  h->g.xcd_remap = xcd_remap;
  if (h->graph) {  // a captured graph holds the old launch shapes
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return WD_OK;
}

int wd_set_precision(wd_handle* h, int mode) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (mode != WD_PREC_FP64 && mode != WD_PREC_MIXED) return fail(h, WD_EINVAL, "precision must be WD_PREC_FP64 or WD_PREC_MIXED");
  if (mode == WD_PREC_MIXED && h->nl.moist) return fail(h, WD_EINVAL, "the mixed form runs dry states only");
  h->nl.mixed = mode == WD_PREC_MIXED;
  // the mixed form runs the column-march kernels (the LDS-tiled kernel reads FP64 tendencies)
  if (h->nl.mixed) h->column_tile = wd::kColumnMarch;
  if (h->graph) {  // a captured graph holds the previous form's kernels
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return WD_OK;
}

int wd_get_precision(const wd_handle* h) { return h ? (h->nl.mixed ? WD_PREC_MIXED : WD_PREC_FP64) : WD_EINVAL; }

int wd_tuned_config(const wd_handle* h, int* column_tile, int* xcd_remap) {
  if (!h || !column_tile || !xcd_remap) return WD_EINVAL;
  *column_tile = h->column_tile;
  *xcd_remap = h->g.xcd_remap;
  return WD_OK;
}

void wd_batch_destroy(wd_batch* b);

int wd_batch_create(wd_handle* const* members, const wd_state* states, int n_members, int branches,
                    wd_batch** out) {
  if (!out) return fail(nullptr, WD_EINVAL, "out is NULL");
  *out = nullptr;
  if (!members || !states || n_members < 1) return fail(nullptr, WD_EINVAL, "need at least one member");
  if (branches < 1) return fail(nullptr, WD_EINVAL, "branches must be >= 1");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  for (int m = 0; m < n_members; ++m) {
    if (!members[m]) return fail(nullptr, WD_EINVAL, "a member handle is NULL");
    if (members[m]->device != dev) return fail(nullptr, WD_EINVAL, "members must live on the current device");
    if (members[m]->prof.mask) return fail(nullptr, WD_EINVAL, "disable profiling on members before batching");
    for (int q = 0; q < m; ++q)
      if (members[q] == members[m]) return fail(nullptr, WD_EINVAL, "a handle appears twice (members share no workspace)");
  }
  wd_batch* b = new (std::nothrow) wd_batch();
  if (!b) return fail(nullptr, WD_ENOMEM, "host allocation failed");
  b->device = dev;
  b->members.assign(members, members + n_members);
  b->states.resize(n_members);
  for (int m = 0; m < n_members; ++m) {
    int rc = check_state(nullptr, &states[m], &b->states[m], members[m]->nl.moist);
    if (rc != WD_OK) {
      delete b;
      return rc;
    }
  }
  const int nb = branches < n_members ? branches : n_members;
  b->streams.assign(nb, nullptr);
  b->events.assign(nb, nullptr);  // [0] fork, [k >= 1] join of branch k
  for (int k = 0; k < nb; ++k) {
    e = hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking);
    if (e != hipSuccess) break;
  }
  for (int k = 0; e == hipSuccess && k < nb; ++k) e = hipEventCreateWithFlags(&b->events[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return hip_fail(nullptr, e, "batch stream/event creation");
  }
  b->branches = nb;
  b->branch_of = assign_branches(b->members, nb);
  int rc = capture_batch(b, b->states, b->branch_of, nb, &b->graph);
  if (rc != WD_OK) {
    wd_batch_destroy(b);
    return rc;
  }
  *out = b;
  return WD_OK;
}

// A grid barrier wait of the persistent step gives up after this many polls
// (a poll is one device-scope load and a short sleep, about a microsecond),
// seconds beyond the longest phase of any ensemble that fits on the GPU.
constexpr unsigned kDefaultSpinLimit = 1u << 22;

int wd_batch_create_persistent(wd_handle* const* members, const wd_state* states, int n_members, wd_batch** out) {
  if (!out) return fail(nullptr, WD_EINVAL, "out is NULL");
  *out = nullptr;
  if (!members || !states || n_members < 1) return fail(nullptr, WD_EINVAL, "need at least one member");
  for (int m = 0; m < n_members; ++m) {
    if (!members[m]) return fail(nullptr, WD_EINVAL, "a member handle is NULL");
    for (int q = 0; q < m; ++q)
      if (members[q] == members[m]) return fail(nullptr, WD_EINVAL, "a handle appears twice (members share no workspace)");
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  for (int m = 0; m < n_members; ++m)
    if (members[m]->device != dev) return fail(nullptr, WD_EINVAL, "members must live on the current device");
  std::vector<wd::PMember> pm(n_members);
  std::vector<int> tiles(n_members);
  for (int m = 0; m < n_members; ++m) {
    wd::DState ds;
    int rc = check_state(nullptr, &states[m], &ds, members[m]->nl.moist);
    if (rc != WD_OK) return rc;
    pm[m] = wd::PMember{members[m]->g, members[m]->nl, ds, members[m]->w};
    tiles[m] = members[m]->column_tile;
  }
  std::vector<int> column_kernel(n_members);
  for (int m = 0; m < n_members; ++m)
    column_kernel[m] = wd::persistent_column_levels(pm[m].g, pm[m].nl, tiles[m]) > 0 ? tiles[m] : wd::kColumnMarch;
  std::vector<wd::PSeg> segs;
  std::vector<wd::PPhase> phases;
  std::string why;
  if (!wd::persistent_program(pm, tiles, &segs, &phases, &why)) return fail(nullptr, WD_EINVAL, why);
  const bool mixed = pm[0].nl.mixed != 0;
  const bool tile = std::any_of(column_kernel.begin(), column_kernel.end(), [](int k) { return k >= 0; });
  // as many workgroups as the GPU holds at once (the grid barrier needs every
  // one resident), a multiple of the 8 XCDs,
  int per_cu = 0, cus = 0;
  e = wd::persistent_blocks_per_cu(mixed, tile, &per_cu);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "occupancy of the persistent step");
  // and no more than the largest phase has virtual blocks: fewer workgroups
  // meet at each barrier, and none of them would find work
  unsigned most = 0;
  for (const wd::PPhase& p : phases) most = p.total > most ? p.total : most;
  const unsigned fit = (unsigned)(per_cu * cus) / wd::kNumXcd * wd::kNumXcd;
  const unsigned need = (most + wd::kNumXcd - 1) / wd::kNumXcd * wd::kNumXcd;
  const unsigned grid = need < fit ? need : fit;
  if (grid == 0) return fail(nullptr, WD_EDEVICE, "the persistent step does not fit on this GPU");
  wd_batch* b = new (std::nothrow) wd_batch();
  if (!b) return fail(nullptr, WD_ENOMEM, "host allocation failed");
  b->device = dev;
  b->persistent = true;
  b->grid = grid;
  b->column_kernel = column_kernel;
  b->members.assign(members, members + n_members);
  b->states.resize(n_members);
  for (int m = 0; m < n_members; ++m) b->states[m] = pm[m].s;
  b->branches = 1;
  b->branch_of.assign(n_members, 0);
  // one allocation: the barrier, the members, the segments, the phases
  const wd::PBarrier bar0{0u, 0u, 0u, kDefaultSpinLimit};
  const size_t o_mem = 256, o_seg = o_mem + slot_bytes((sizeof(wd::PMember) * pm.size() + 7) / 8),
               o_ph = o_seg + slot_bytes((sizeof(wd::PSeg) * segs.size() + 7) / 8),
               total = o_ph + slot_bytes((sizeof(wd::PPhase) * phases.size() + 7) / 8);
  e = hipMalloc(&b->pbuf, total);
  char* base = static_cast<char*>(b->pbuf);
  if (e == hipSuccess) e = hipMemcpy(base, &bar0, sizeof(bar0), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(base + o_mem, pm.data(), sizeof(wd::PMember) * pm.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(base + o_seg, segs.data(), sizeof(wd::PSeg) * segs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(base + o_ph, phases.data(), sizeof(wd::PPhase) * phases.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return hip_fail(nullptr, e, "persistent step tables");
  }
  b->bar = reinterpret_cast<wd::PBarrier*>(base);
  // capture the one launch per large step
  b->streams.assign(1, nullptr);
  e = hipStreamCreateWithFlags(&b->streams[0], hipStreamNonBlocking);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return hip_fail(nullptr, e, "batch stream creation");
  }
  hipStream_t s0 = b->streams[0];
  e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return hip_fail(nullptr, e, "hipStreamBeginCapture");
  }
  e = wd::launch_persistent_step(mixed, tile, grid, reinterpret_cast<const wd::PMember*>(base + o_mem),
                                 reinterpret_cast<const wd::PSeg*>(base + o_seg),
                                 reinterpret_cast<const wd::PPhase*>(base + o_ph), (int)phases.size(), b->bar, s0);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s0, &graph);
  if (e == hipSuccess) e = ec;
  if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return hip_fail(nullptr, e, "capturing the persistent step");
  }
  *out = b;
  return WD_OK;
}

int wd_batch_status(wd_batch* b, void* stream) {
  if (!b) return fail(nullptr, WD_EINVAL, "batch is NULL");
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    b->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
    return WD_EDEVICE;
  }
  if (!b->persistent) return WD_OK;
  wd::PBarrier bar{};
  e = hipMemcpy(&bar, b->bar, sizeof(bar), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    b->err = std::string("reading the grid barrier: ") + hipGetErrorString(e);
    return WD_EDEVICE;
  }
  if (!bar.abort) return WD_OK;
  // nothing runs now: clear the flag and the arrivals for the next launches
  bar.count = 0;
  bar.abort = 0;
  e = hipMemcpy(b->bar, &bar, sizeof(bar), hipMemcpyHostToDevice);
  b->err = "a grid barrier wait of the persistent step ran out (not every workgroup was resident at once, "
           "e.g. beside other work on the GPU): the large steps since the last check did not complete";
  if (e != hipSuccess) b->err += std::string("; clearing the flag: ") + hipGetErrorString(e);
  return WD_EDEVICE;
}

int wd_batch_set_spin_limit(wd_batch* b, unsigned polls) {
  if (!b) return fail(nullptr, WD_EINVAL, "batch is NULL");
  if (!b->persistent) {
    b->err = "only a persistent batch has a grid barrier";
    return WD_EINVAL;
  }
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(&b->bar->spin_limit, &polls, sizeof(polls), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    b->err = std::string("setting the spin limit: ") + hipGetErrorString(e);
    return WD_EDEVICE;
  }
  return WD_OK;
}

int wd_batch_workgroups(const wd_batch* b) { return b ? (int)b->grid : WD_EINVAL; }

int wd_batch_column_kernel(const wd_batch* b, int member, int* column_tile) {
  if (!b || !column_tile || member < 0 || member >= (int)b->members.size()) return WD_EINVAL;
  *column_tile = b->persistent ? b->column_kernel[member] : b->members[member]->column_tile;
  return WD_OK;
}

int wd_step_batched(wd_batch* b, int n_large_steps, void* stream) {
  if (!b) return fail(nullptr, WD_EINVAL, "batch is NULL");
  if (n_large_steps < 0) {
    b->err = "n_large_steps must be >= 0";
    return WD_EINVAL;
  }
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != b->device) {
    b->err = "the current HIP device is not the batch's device (" + std::to_string(b->device) + ")";
    return WD_EINVAL;
  }
  for (int i = 0; i < n_large_steps; ++i) {
    hipError_t e = hipGraphLaunch(b->graph, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
      b->err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
      return WD_EDEVICE;
    }
  }
  return WD_OK;
}

int wd_batch_branch(const wd_batch* b, int member) {
  if (!b || member < 0 || member >= (int)b->members.size()) return WD_EINVAL;
  return b->branch_of[member];
}

int wd_batch_branches(const wd_batch* b) { return b ? b->branches : WD_EINVAL; }

const char* wd_batch_last_error(const wd_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

void wd_batch_destroy(wd_batch* b) {
  if (!b) return;
  int cur = 0;
  if (hipGetDevice(&cur) == hipSuccess && cur != b->device) hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  if (b->graph) hipGraphExecDestroy(b->graph);
  if (b->pbuf) hipFree(b->pbuf);
  for (hipEvent_t x : b->events)
    if (x) hipEventDestroy(x);
  for (hipStream_t x : b->streams)
    if (x) hipStreamDestroy(x);
  if (cur != b->device) hipSetDevice(cur);
  delete b;
}

int wd_profile_enable(wd_handle* h, unsigned long long kernel_mask) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  h->prof.mask = kernel_mask;
  h->prof.recs.clear();
  h->prof.used = 0;
  h->prof.failed = false;
  return WD_OK;
}

int wd_profile_read(wd_handle* h, double* total_ms, long long* launches, int n) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (n < 0 || (n > 0 && (!total_ms || !launches))) return fail(h, WD_EINVAL, "bad output arrays");
  for (int i = 0; i < n; ++i) {
    total_ms[i] = 0.0;
    launches[i] = 0;
  }
  wd::Prof& p = h->prof;
  if (p.failed) return fail(h, WD_EDEVICE, "event pool allocation failed while profiling");
  for (const auto& r : p.recs) {
    hipEvent_t a = p.pool[r.second], b = p.pool[r.second + 1];
    hipError_t e = hipEventSynchronize(b);
    if (e != hipSuccess) return hip_fail(h, e, "hipEventSynchronize");
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess) return hip_fail(h, e, "hipEventElapsedTime");
    if (r.first < n) {
      total_ms[r.first] += (double)ms;
      launches[r.first] += 1;
    }
  }
  p.recs.clear();
  p.used = 0;
  return wd::KID_COUNT;
}

int wd_kernel_count(void) { return wd::KID_COUNT; }

const char* wd_kernel_name(int id) { return wd::kernel_name(id); }

size_t wd_workspace_bytes(const wd_handle* h) { return h ? h->work_bytes + h->metric_bytes : 0; }

const char* wd_last_error(const wd_handle* h) {
  if (h) return h->err.c_str();
  return g_create_error.c_str();
}

void wd_destroy(wd_handle* h) {
  if (!h) return;
  int cur = 0;
  if (hipGetDevice(&cur) == hipSuccess && cur != h->device) hipSetDevice(h->device);
  // Work queued on any stream may still use the workspace.
  (void)hipDeviceSynchronize();
  if (h->graph) hipGraphExecDestroy(h->graph);
  if (h->capture_stream) hipStreamDestroy(h->capture_stream);
  if (h->work_buf) hipFree(h->work_buf);
  if (h->diff_buf) hipFree(h->diff_buf);
  if (h->metric_buf) hipFree(h->metric_buf);
  wd::adjoint_tape_destroy(h->tape);
  if (cur != h->device) hipSetDevice(cur);
  delete h;
}

int wd_step_adjoint(wd_handle* h, const wd_state* start, wd_state* bar, void* stream) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (h->nl.moist) return fail(h, WD_EINVAL, "the adjoint does not cover the moisture (moist = 1)");
  wd::DState ds, db;
  int rc = check_state(h, start, &ds);
  if (rc == WD_OK) rc = check_state(h, bar, &db);
  if (rc == WD_OK) rc = check_device(h);
  if (rc != WD_OK) return rc;
  if (h->nl.mixed) return fail(h, WD_EINVAL, "the adjoint is built for the FP64 form (WD_PREC_FP64) only");
  if (h->nl.diff_6th_opt != 0) return fail(h, WD_EINVAL, "the adjoint does not cover the 6th-order filter");
  if (h->nl.smag) return fail(h, WD_EINVAL, "the adjoint does not cover the Smagorinsky diffusion");
  if (h->nl.const_nu > 0.0) return fail(h, WD_EINVAL, "the adjoint does not cover the constant-K diffusion");
  if (h->nl.scalar_adv_opt != 0) return fail(h, WD_EINVAL, "the adjoint does not cover the theta limiter");
  if (h->g.hypsometric_opt == 2) return fail(h, WD_EINVAL, "the adjoint does not cover hypsometric_opt=2");
  if (h->g.rot) return fail(h, WD_EINVAL, "the adjoint does not cover the Coriolis terms (f or e nonzero)");
  // Switched off: the first GPU run of this path ended in a memory access
  // fault whose cause is not yet found (its kernel bodies are bounds-checked
  // on the host).  It stays off the GPU until the cause is shown.
  return fail(h, WD_EINVAL, "wd_step_adjoint is switched off: its first GPU run faulted, cause not yet found");
  const double* a[] = {ds.u, ds.v, ds.w, ds.theta, ds.p_total, ds.p_pert, ds.ph_total, ds.ph_pert,
                       ds.mu_total, ds.mu_pert, ds.mudf};
  const double* b[] = {db.u, db.v, db.w, db.theta, db.p_total, db.p_pert, db.ph_total, db.ph_pert,
                       db.mu_total, db.mu_pert, db.mudf};
  for (const double* x : a)
    for (const double* y : b)
      if (x == y) return fail(h, WD_EINVAL, "a start array and a bar array are the same buffer");
  if (!h->tape) {
    hipError_t e = wd::adjoint_tape_create(h->g, h->nl, &h->tape, &h->tape_bytes);
    if (e != hipSuccess)
      return fail(h, WD_ENOMEM, std::string("allocating the adjoint tape: ") + hipGetErrorString(e));
  }
  hipError_t e = wd::adjoint_large_step(h->g, h->nl, h->w, h->tape, ds, db, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(h, e, "adjoint step");
  return WD_OK;
}

size_t wd_adjoint_tape_bytes(const wd_handle* h) { return h ? h->tape_bytes : 0; }

int wd_step_noalloc(const wd_grid* grid, const wd_namelist* nl, const wd_state* st, int n_large_steps, void* stream) {
  bool rot = false;
  int rc = validate_create(grid, nl, &rot);
  if (rc != WD_OK) return rc;
  if (n_large_steps < 0) return fail(nullptr, WD_EINVAL, "n_large_steps must be >= 0");
  wd::DState ds;
  rc = check_state(nullptr, st, &ds, nl->moist);
  if (rc != WD_OK) return rc;
  wd_handle h;  // on the stack: its device buffers all come from `scratch`
  init_params(grid, nl, rot, h.g, h.nl);
  h.column_tile = wd::kColumnMarch;
  // the layout in `scratch`: the metrics, then the workspace (base, then the
  // optional terms' arrays), each 256-byte aligned
  const size_t n1 = grid->nz, n1f = (size_t)grid->nz + 1;
  const size_t c = (size_t)grid->ny * grid->nx, cu = (size_t)grid->ny * (grid->nx + 1),
               cv = (size_t)(grid->ny + 1) * grid->nx;
  struct Metric {
    const double* src;  // host array, or NULL: filled with `value`
    size_t n;
    double value;
    const double** dst;
  };
  wd::DGrid& g = h.g;
  const Metric metrics[] = {
      {grid->c1h, n1, 0.0, &g.c1h},     {grid->c2h, n1, 0.0, &g.c2h},     {grid->c1f, n1f, 0.0, &g.c1f},
      {grid->c2f, n1f, 0.0, &g.c2f},    {grid->dnw, n1, 0.0, &g.dnw},     {grid->rdnw, n1, 0.0, &g.rdnw},
      {grid->dn, n1, 0.0, &g.dn},       {grid->rdn, n1, 0.0, &g.rdn},     {grid->fnm, n1, 0.0, &g.fnm},
      {grid->fnp, n1, 0.0, &g.fnp},     {grid->msftx, c, 0.0, &g.msftx},  {grid->msfty, c, 0.0, &g.msfty},
      {grid->msfux, cu, 0.0, &g.msfux}, {grid->msfuy, cu, 0.0, &g.msfuy}, {grid->msfvx, cv, 0.0, &g.msfvx},
      {grid->msfvy, cv, 0.0, &g.msfvy}, {grid->c3h, n1, 1.0, &g.c3h},     {grid->c4h, n1, 0.0, &g.c4h},
      {grid->c3f, n1f, 1.0, &g.c3f},    {grid->c4f, n1f, 0.0, &g.c4f},    {grid->f, c, 0.0, &g.fcor},
      {grid->e, c, 0.0, &g.ecor},       {grid->sina, c, 0.0, &g.sina},    {grid->cosa, c, 1.0, &g.cosa},
  };
  const int n_metrics = rot ? 24 : 20;  // the rotation arrays only when the grid rotates
  WsSlot ws[2 * kMaxWsSlots];
  const int nbase = workspace_slots(g, h.nl, false, h.w, ws);
  const int nws = nbase + workspace_slots(g, h.nl, true, h.w, ws + nbase);
  size_t total = 0;
  for (int q = 0; q < n_metrics; ++q) total += slot_bytes(metrics[q].n);
  for (int q = 0; q < nws; ++q) total += slot_bytes(ws[q].n);
  if (total > kScratchBytes)
    return fail(nullptr, WD_EINVAL,
                "wd_step_noalloc: this grid and namelist need " + std::to_string(total) +
                    " bytes of device scratch, above its fixed " + std::to_string(kScratchBytes) +
                    " (use wd_create and wd_step)");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  double* base = nullptr;
  rc = scratch_base(dev, &base);
  if (rc != WD_OK) return rc;
  h.device = dev;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  char* cur = reinterpret_cast<char*>(base);
  for (int q = 0; q < n_metrics; ++q) {
    const Metric& m = metrics[q];
    double* d = reinterpret_cast<double*>(cur);
    *m.dst = d;
    if (m.src)
      e = hipMemcpyAsync(d, m.src, m.n * sizeof(double), hipMemcpyHostToDevice, s);
    else {
      k_fill<<<(unsigned)std::min<size_t>((m.n + 255) / 256, 1024), 256, 0, s>>>(d, m.n, m.value);
      e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_fail(nullptr, e, "wd_step_noalloc: metrics to the device");
    cur += slot_bytes(m.n);
  }
  // the caller's host metric arrays are read by the copies above: wait for them
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipStreamSynchronize");
  // the workspace starts at zero, as wd_create leaves it
  char* ws_begin = cur;
  for (int q = 0; q < nws; ++q) {
    *ws[q].p = reinterpret_cast<double*>(cur);
    cur += slot_bytes(ws[q].n);
  }
  const size_t ws_doubles = (size_t)(cur - ws_begin) / sizeof(double);
  k_fill<<<(unsigned)std::min<size_t>((ws_doubles + 255) / 256, 4096), 256, 0, s>>>(
      reinterpret_cast<double*>(ws_begin), ws_doubles, 0.0);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(nullptr, e, "wd_step_noalloc: clearing the workspace");
  rc = enqueue_steps(&h, ds, n_large_steps, s);
  if (rc != WD_OK) return fail(nullptr, rc, h.err);
  return WD_OK;
}

}  // extern "C"
