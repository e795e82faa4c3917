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
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "wrfdyn.h"
#include "wrfdyn_kernels.h"

struct wd_handle {
  wd::DGrid g{};
  wd::DNamelist nl{};
  wd::DWork w{};
  void* metric_buf = nullptr;
  void* work_buf = nullptr;
  size_t work_bytes = 0;
  size_t metric_bytes = 0;
  int device = 0;
  int column_tile = 0;
  std::string err;
  wd::Prof prof;
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

}  // namespace

extern "C" {

int wd_create(const wd_grid* grid, const wd_namelist* nl, wd_handle** out) {
  if (!out) return fail(nullptr, WD_EINVAL, "out is NULL");
  *out = nullptr;
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
  host.reserve(8 * n1 + 2 * n1f + 2 * c + 2 * cu + 2 * cv);
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
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.dx = grid->dx;
  g.dy = grid->dy;
  g.rdx = 1.0 / grid->dx;
  g.rdy = 1.0 / grid->dy;
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
  g.cf1 = grid->cf1;
  g.cf2 = grid->cf2;
  g.cf3 = grid->cf3;
  // WRF top-face extrapolation weights cfn/cfn1 (core/rk_addtend_dry.py:305-312).
  double dn_top = grid->dn[nz - 1];
  double dn_safe = std::fabs(dn_top) > 1.0e-30 ? dn_top : 1.0;
  g.cfn = (0.5 * grid->dnw[nz - 1] + grid->dn[nz - 1]) / dn_safe;
  g.cfn1 = -0.5 * grid->dnw[nz - 1] / dn_safe;

  h->nl.dt = nl->dt;
  h->nl.ns = nl->acoustic_substeps;
  h->nl.epssm = nl->epssm;
  h->nl.top_lid = nl->top_lid;
  h->nl.w_damping = nl->w_damping;
  h->nl.damp_opt = nl->damp_opt;
  h->nl.dampcoef = nl->dampcoef;
  h->nl.zdamp = nl->zdamp;

  // ---- workspace ----
  const size_t M = (size_t)nz * ny * nx, Wn = (size_t)(nz + 1) * ny * nx;
  const size_t U = (size_t)nz * ny * (nx + 1), V = (size_t)nz * (ny + 1) * nx;
  struct Slot {
    double** p;
    size_t n;
  };
  wd::DWork& w = h->w;
  const Slot slots[] = {
      {&w.u0, U},       {&w.v0, V},        {&w.w0, Wn},      {&w.th0, M},      {&w.ph0, Wn},
      {&w.mu0, c},      {&w.ru, M},        {&w.rv, M},       {&w.rom, Wn},     {&w.ut, U},
      {&w.vt, V},       {&w.wt, Wn},       {&w.tht, M},      {&w.alt_raw, M},  {&w.al_ls, M},
      {&w.php_ls, M},   {&w.mub, c},       {&w.mut, c},      {&w.muts0, c},    {&w.mu_work0, c},
      {&w.mu, c},       {&w.muts, c},      {&w.muave, c},    {&w.muu, cu},     {&w.muus, cu},
      {&w.muv, cv},     {&w.muvs, cv},     {&w.u_work, U},   {&w.v_work, V},   {&w.w_work, Wn},
      {&w.th_work, M},  {&w.ph_work, Wn},  {&w.c2a, M},      {&w.alt, M},      {&w.pb, M},
      {&w.ph1t, Wn},     {&w.php, M},       {&w.p, M},        {&w.al, M},       {&w.pm1, M},
      {&w.rw_tend, Wn}, {&w.ph_tend, Wn},  {&w.ca, Wn},      {&w.calpha, Wn},  {&w.cgamma, Wn},
      {&w.ww, Wn},      {&w.t2ave, M},     {&w.rhs, Wn},     {&w.wfwd, Wn},
  };
  // 256-byte aligned sub-buffers
  size_t total = 0;
  for (const Slot& s : slots) total += (s.n * sizeof(double) + 255) / 256 * 256;
  e = hipMalloc(&h->work_buf, total);
  if (e != hipSuccess) {
    hipFree(h->metric_buf);
    delete h;
    return fail(nullptr, WD_ENOMEM, std::string("hipMalloc(workspace): ") + hipGetErrorString(e));
  }
  h->work_bytes = total;
  char* base = static_cast<char*>(h->work_buf);
  size_t off = 0;
  for (const Slot& s : slots) {
    *s.p = reinterpret_cast<double*>(base + off);
    off += (s.n * sizeof(double) + 255) / 256 * 256;
  }
  e = hipMemset(h->work_buf, 0, total);
  if (e != hipSuccess) {
    hipFree(h->work_buf);
    hipFree(h->metric_buf);
    delete h;
    return hip_fail(nullptr, e, "hipMemset(workspace)");
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
  // XCD-contiguous tile order (default on); WRFDYN_XCD_REMAP=0 keeps the
  // hardware dispatch order, for measurement.
  h->g.xcd_remap = 1;
  if (const char* env = std::getenv("WRFDYN_XCD_REMAP")) h->g.xcd_remap = std::atoi(env) != 0;
  *out = h;
  return WD_OK;
}

int wd_step(wd_handle* h, const wd_state* st, int n_large_steps, void* stream) {
  if (!h) return fail(nullptr, WD_EINVAL, "handle is NULL");
  if (!st) return fail(h, WD_EINVAL, "state is NULL");
  if (n_large_steps < 0) return fail(h, WD_EINVAL, "n_large_steps must be >= 0");
  double* ptrs[] = {st->u, st->v, st->w, st->theta, st->p_total, st->p_pert, st->ph_total,
                    st->ph_pert, st->mu_total, st->mu_pert, st->mudf};
  for (double* p : ptrs)
    if (!p) return fail(h, WD_EINVAL, "a state pointer is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  wd::DState ds{st->u, st->v, st->w, st->theta, st->p_total, st->p_pert,
                st->ph_total, st->ph_pert, st->mu_total, st->mu_pert, st->mudf};
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
    }
  }
  return WD_OK;
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
  if (h->work_buf) hipFree(h->work_buf);
  if (h->metric_buf) hipFree(h->metric_buf);
  if (cur != h->device) hipSetDevice(cur);
  delete h;
}

}  // extern "C"
