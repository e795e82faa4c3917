// Adjoint of one plain-FP64 large step on the GPU: the vector-Jacobian
// product of _rk_scan_step (src/gpuwrf/runtime/operational_mode.py:3319-3534)
// with respect to every array of the state it starts from.
//
// Forward recompute with a tape.  From the step's starting state, the step
// is rerun with the product kernels (stage setup, advance_uv, the
// column-march advance_mu_t and advance_w + calc_p_rho, finish), and the
// arrays each adjoint piece reads are copied out at every piece boundary:
// per stage the state and the stage constants the setup computes, and at
// every boundary the carried acoustic arrays.  Two places where the forward
// kernels take a value from a flag rather than from the array (t_2ave = 0
// and pm1 = p in a stage's first substep) are written into the copies, so
// each copy holds the values the reference's piece reads.
//
// Reverse.  Stage by stage from the last: finish, the substeps in reverse
// (calc_p_rho, advance_w, advance_mu_t + theta, advance_uv), then the stage
// setup in its phases; finally the step origin.  The bodies are those of
// wrfdyn_adjoint.h, one thread per column, face column or point; bars that
// neighbouring threads share are updated with atomic adds, so the result
// is the same to rounding, not bit for bit, from run to run.
#include <vector>

#include "wrfdyn_adjoint.h"
#include "wrfdyn_kernels.h"

namespace wd {

using namespace adj;

namespace {

struct AtomicAdd {
  __device__ void operator()(double* p, double v) const { atomicAdd(p, v); }
};

enum Ext : int { E_M, E_W, E_U, E_V, E_C, E_CU, E_CV };
// extents of the AdjArr arrays, in their order
constexpr int kExtent[A_count] = {
    E_U, E_V, E_W, E_M, E_M, E_M, E_W, E_W, E_C, E_C, E_C,
    E_U, E_V, E_W, E_M, E_W, E_C, E_M, E_M, E_W, E_M, E_U, E_V, E_W, E_M,
    E_M, E_M, E_M, E_C, E_C, E_C, E_C, E_CU, E_CU, E_CV, E_CV,
    E_U, E_V, E_W, E_M, E_W, E_M, E_M, E_M, E_W, E_M, E_M, E_M,
    E_M, E_W, E_W, E_W, E_W, E_W, E_C, E_C, E_C, E_W, E_M,
    E_M, E_C, E_W, E_W, E_W, E_W};

size_t ext_len(const DGrid& g, int e) {
  const size_t nx = g.nx, ny = g.ny, nz = g.nz;
  switch (e) {
    case E_M: return nz * ny * nx;
    case E_W: return (nz + 1) * ny * nx;
    case E_U: return nz * ny * (nx + 1);
    case E_V: return nz * (ny + 1) * nx;
    case E_C: return ny * nx;
    case E_CU: return ny * (nx + 1);
    default: return (ny + 1) * nx;
  }
}

// The state fields kept per stage (mudf changes inside a stage and is carried).
constexpr int kStageState[] = {A_u, A_v, A_w, A_theta, A_p_total, A_p_pert, A_ph_total, A_ph_pert, A_mu_total,
                               A_mu_pert};
// Work arrays that the setup computes and the substeps and finish only read.
constexpr int kStageConst[] = {A_u0, A_v0, A_w0, A_th0, A_ph0, A_mu0, A_ru, A_rv, A_rom, A_ut, A_vt, A_wt,
                               A_tht, A_alt_raw, A_al_ls, A_php_ls, A_mub, A_mut, A_muts0, A_mu_work0, A_muu,
                               A_muus, A_muv, A_muvs, A_c2a, A_alt, A_pb, A_php, A_rw_tend, A_ph_tend, A_ca,
                               A_calpha, A_cgamma, A_phb};
// Arrays that change from substep to substep.
constexpr int kCarried[] = {A_u_work, A_v_work, A_w_work, A_th_work, A_ph_work, A_p, A_al, A_pm1,
                            A_mu, A_muts, A_muave, A_ww, A_t2ave, A_mudf};
constexpr int kNStageState = sizeof(kStageState) / sizeof(int);
constexpr int kNStageConst = sizeof(kStageConst) / sizeof(int);
constexpr int kNCarried = sizeof(kCarried) / sizeof(int);

double* dwork_ptr(const DWork& w, int a) {
  switch (a) {
    case A_u0: return w.u0;
    case A_v0: return w.v0;
    case A_w0: return w.w0;
    case A_th0: return w.th0;
    case A_ph0: return w.ph0;
    case A_mu0: return w.mu0;
    case A_ru: return w.ru;
    case A_rv: return w.rv;
    case A_rom: return w.rom;
    case A_ut: return w.ut;
    case A_vt: return w.vt;
    case A_wt: return w.wt;
    case A_tht: return w.tht;
    case A_alt_raw: return w.alt_raw;
    case A_al_ls: return w.al_ls;
    case A_php_ls: return w.php_ls;
    case A_mub: return w.mub;
    case A_mut: return w.mut;
    case A_muts0: return w.muts0;
    case A_mu_work0: return w.mu_work0;
    case A_muu: return w.muu;
    case A_muus: return w.muus;
    case A_muv: return w.muv;
    case A_muvs: return w.muvs;
    case A_u_work: return w.u_work;
    case A_v_work: return w.v_work;
    case A_w_work: return w.w_work;
    case A_th_work: return w.th_work;
    case A_ph_work: return w.ph_work;
    case A_c2a: return w.c2a;
    case A_alt: return w.alt;
    case A_pb: return w.pb;
    case A_php: return w.php;
    case A_p: return w.p;
    case A_al: return w.al;
    case A_pm1: return w.pm1;
    case A_rw_tend: return w.rw_tend;
    case A_ph_tend: return w.ph_tend;
    case A_ca: return w.ca;
    case A_calpha: return w.calpha;
    case A_cgamma: return w.cgamma;
    case A_mu: return w.mu;
    case A_muts: return w.muts;
    case A_muave: return w.muave;
    case A_ww: return w.ww;
    case A_t2ave: return w.t2ave;
    default: return nullptr;
  }
}

double* state_ptr(const DState& s, int a) {
  switch (a) {
    case A_u: return s.u;
    case A_v: return s.v;
    case A_w: return s.w;
    case A_theta: return s.theta;
    case A_p_total: return s.p_total;
    case A_p_pert: return s.p_pert;
    case A_ph_total: return s.ph_total;
    case A_ph_pert: return s.ph_pert;
    case A_mu_total: return s.mu_total;
    case A_mu_pert: return s.mu_pert;
    case A_mudf: return s.mudf;
    default: return nullptr;
  }
}

constexpr unsigned kBs = 256;
unsigned nblocks(size_t n) { return (unsigned)((n + kBs - 1) / kBs); }

// ---- kernels: one thread per column / face column / point ----
__global__ void __launch_bounds__(kBs) k_adj_phb(size_t n, const double* pht, const double* php, double* out) {
  const size_t q = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (q < n) out[q] = pht[q] - php[q];
}
__global__ void __launch_bounds__(kBs) k_adj_origin(View V, int src, int dst, size_t n) {
  const size_t q = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (q < n) adj_origin_elem(V, src, dst, q);
}
__global__ void __launch_bounds__(kBs) k_adj_finish_col(DGrid g, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_finish_column(g, V, (int)(c / g.nx), (int)(c % g.nx));
}
// u faces (ny x (nx+1)) then v faces ((ny+1) x nx)
__device__ inline bool face_of(const DGrid& g, size_t f, bool* ydir, int* j, int* i) {
  const size_t nu = (size_t)g.ny * (g.nx + 1);
  if (f < nu) {
    *ydir = false;
    *j = (int)(f / (g.nx + 1));
    *i = (int)(f % (g.nx + 1));
    return true;
  }
  f -= nu;
  if (f >= (size_t)(g.ny + 1) * g.nx) return false;
  *ydir = true;
  *j = (int)(f / g.nx);
  *i = (int)(f % g.nx);
  return true;
}
size_t n_faces(const DGrid& g) { return (size_t)g.ny * (g.nx + 1) + (size_t)(g.ny + 1) * g.nx; }

__global__ void __launch_bounds__(kBs) k_adj_finish_face(DGrid g, View V) {
  bool yd;
  int j, i;
  if (face_of(g, (size_t)blockIdx.x * kBs + threadIdx.x, &yd, &j, &i)) adj_finish_face(g, V, yd, j, i);
}
__global__ void __launch_bounds__(kBs) k_adj_p(DGrid g, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_p_column(g, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_w(DGrid g, DNamelist nl, double dts, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_w_column(g, nl, dts, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_mu_theta(DGrid g, DNamelist nl, double dts, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_mu_theta_column(g, nl, dts, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_uv(DGrid g, double dts, View V) {
  bool yd;
  int j, i;
  if (face_of(g, (size_t)blockIdx.x * kBs + threadIdx.x, &yd, &j, &i)) adj_uv_face(g, dts, V, yd, j, i, AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_s1(DGrid g, DNamelist nl, double dts, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_setup_s1_column(g, nl, dts, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_s2(DGrid g, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_setup_s2_column(g, V, (int)(c / g.nx), (int)(c % g.nx));
}
__global__ void __launch_bounds__(kBs) k_adj_s3(DGrid g, View V) {
  bool yd;
  int j, i;
  if (face_of(g, (size_t)blockIdx.x * kBs + threadIdx.x, &yd, &j, &i)) adj_setup_s3_face(g, V, yd, j, i);
}
__global__ void __launch_bounds__(kBs) k_adj_s4(DGrid g, int rk_step, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_setup_s4_column(g, rk_step, V, (int)(c / g.nx), (int)(c % g.nx));
}
__global__ void __launch_bounds__(kBs) k_adj_s4_clear(DGrid g, View V) {
  bool yd;
  int j, i;
  if (face_of(g, (size_t)blockIdx.x * kBs + threadIdx.x, &yd, &j, &i)) adj_setup_s4_clear(g, V, yd, j, i);
}
__global__ void __launch_bounds__(kBs) k_adj_pgf(DGrid g, View V) {
  bool yd;
  int j, i;
  if (face_of(g, (size_t)blockIdx.x * kBs + threadIdx.x, &yd, &j, &i)) adj_pgf_face(g, V, yd, j, i, AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_s6(DGrid g, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_setup_s6_column(g, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
// periodic tendency copies: u rows (nz x ny) then v columns (nz x nx)
__global__ void __launch_bounds__(kBs) k_adj_adv_copy(DGrid g, View V) {
  const size_t t = (size_t)blockIdx.x * kBs + threadIdx.x;
  const size_t nu = (size_t)g.nz * g.ny, nv = (size_t)g.nz * g.nx;
  if (t < nu) adj_advect_copy(g, V, false, (int)(t / g.ny), (int)(t % g.ny));
  else if (t < nu + nv) adj_advect_copy(g, V, true, (int)((t - nu) / g.nx), (int)((t - nu) % g.nx));
}
__global__ void __launch_bounds__(kBs) k_adj_adv_h(DGrid g, View V, int kind, int nlev) {
  const size_t t = (size_t)blockIdx.x * kBs + threadIdx.x;
  const size_t plane = (size_t)g.ny * g.nx;
  if (t >= (size_t)nlev * plane) return;
  const int k = (int)(t / plane);
  const size_t c = t % plane;
  adj_advect_h_point(g, V, kind, k, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_adv_v(DGrid g, View V, int kind) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_advect_v_column(g, V, kind, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_couple_col(DGrid g, View V) {
  const size_t c = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (c < (size_t)g.ny * g.nx) adj_couple_column(g, V, (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}
__global__ void __launch_bounds__(kBs) k_adj_couple_pt(DGrid g, View V) {
  const size_t t = (size_t)blockIdx.x * kBs + threadIdx.x;
  const size_t plane = (size_t)g.ny * g.nx;
  if (t >= (size_t)g.nz * plane) return;
  const size_t c = t % plane;
  adj_couple_point(g, V, (int)(t / plane), (int)(c / g.nx), (int)(c % g.nx), AtomicAdd{});
}

int n_boundaries(const DNamelist& nl, int* nsub) {
  const int ns = nl.ns;
  nsub[0] = 1;
  nsub[1] = ns / 2 > 1 ? ns / 2 : 1;
  nsub[2] = ns;
  return 3 + 3 * (nsub[0] + nsub[1] + nsub[2]);
}

}  // namespace

struct AdjTape {
  void* buf = nullptr;
  size_t bytes = 0;
  DState scratch{};                                   // forward recompute state
  std::vector<std::vector<double*>> stage_state;      // [3][kNStageState]
  std::vector<std::vector<double*>> stage_const;      // [3][kNStageConst]
  std::vector<std::vector<double*>> carried;          // [boundary][kNCarried]
  double* bar[A_count] = {};
  double* scr = nullptr;                              // kScrSlots planes of (nz+1, ny, nx)
  double* dummy = nullptr;                            // a zero array of the largest extent
};

hipError_t adjoint_tape_create(const DGrid& g, const DNamelist& nl, AdjTape** out, size_t* bytes) {
  *out = nullptr;
  int nsub[3];
  const int nb = n_boundaries(nl, nsub);
  AdjTape* t = new (std::nothrow) AdjTape();
  if (!t) return hipErrorOutOfMemory;
  struct Want {
    double** p;
    size_t n;
  };
  std::vector<Want> want;
  const size_t wlen = ext_len(g, E_W), big = std::max(ext_len(g, E_U), std::max(ext_len(g, E_V), wlen));
  for (int a = 0; a < kNStageState + 1; ++a) {
    const int arr = a < kNStageState ? kStageState[a] : A_mudf;
    double** slot = &t->scratch.u;
    switch (arr) {
      case A_u: slot = &t->scratch.u; break;
      case A_v: slot = &t->scratch.v; break;
      case A_w: slot = &t->scratch.w; break;
      case A_theta: slot = &t->scratch.theta; break;
      case A_p_total: slot = &t->scratch.p_total; break;
      case A_p_pert: slot = &t->scratch.p_pert; break;
      case A_ph_total: slot = &t->scratch.ph_total; break;
      case A_ph_pert: slot = &t->scratch.ph_pert; break;
      case A_mu_total: slot = &t->scratch.mu_total; break;
      case A_mu_pert: slot = &t->scratch.mu_pert; break;
      default: slot = &t->scratch.mudf; break;
    }
    want.push_back({slot, ext_len(g, kExtent[arr])});
  }
  t->stage_state.assign(3, std::vector<double*>(kNStageState));
  t->stage_const.assign(3, std::vector<double*>(kNStageConst));
  t->carried.assign(nb, std::vector<double*>(kNCarried));
  for (int s = 0; s < 3; ++s) {
    for (int a = 0; a < kNStageState; ++a) want.push_back({&t->stage_state[s][a], ext_len(g, kExtent[kStageState[a]])});
    for (int a = 0; a < kNStageConst; ++a) want.push_back({&t->stage_const[s][a], ext_len(g, kExtent[kStageConst[a]])});
  }
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < kNCarried; ++a) want.push_back({&t->carried[b][a], ext_len(g, kExtent[kCarried[a]])});
  for (int a = 0; a < A_count; ++a) want.push_back({&t->bar[a], ext_len(g, kExtent[a])});
  want.push_back({&t->scr, (size_t)kScrSlots * wlen});
  want.push_back({&t->dummy, big});
  size_t total = 0;
  for (const Want& x : want) total += (x.n * sizeof(double) + 255) / 256 * 256;
  hipError_t e = hipMalloc(&t->buf, total);
  if (e != hipSuccess) {
    delete t;
    return e;
  }
  e = hipMemset(t->buf, 0, total);
  if (e != hipSuccess) {
    (void)hipFree(t->buf);
    delete t;
    return e;
  }
  t->bytes = total;
  char* base = static_cast<char*>(t->buf);
  size_t off = 0;
  for (const Want& x : want) {
    *x.p = reinterpret_cast<double*>(base + off);
    off += (x.n * sizeof(double) + 255) / 256 * 256;
  }
  *out = t;
  *bytes = total;
  return hipSuccess;
}

void adjoint_tape_destroy(AdjTape* t) {
  if (!t) return;
  if (t->buf) (void)hipFree(t->buf);
  delete t;
}

namespace {

hipError_t copy_arr(const DGrid& g, int arr, double* dst, const double* src, hipStream_t st) {
  return hipMemcpyAsync(dst, src, ext_len(g, kExtent[arr]) * sizeof(double), hipMemcpyDeviceToDevice, st);
}

// Copy the carried arrays into boundary b; first: the stage's first substep
// has not yet run advance_w, so t_2ave is 0 and pm1 is the step-0 p.
hipError_t snap_carried(const DGrid& g, const DWork& w, const DState& s, AdjTape* t, int b, bool first,
                        hipStream_t st) {
  for (int a = 0; a < kNCarried; ++a) {
    const int arr = kCarried[a];
    const double* src = arr == A_mudf ? s.mudf : dwork_ptr(w, arr);
    hipError_t e;
    if (first && arr == A_t2ave)
      e = hipMemsetAsync(t->carried[b][a], 0, ext_len(g, E_M) * sizeof(double), st);
    else if (first && arr == A_pm1)
      e = copy_arr(g, arr, t->carried[b][a], w.p, st);
    else
      e = copy_arr(g, arr, t->carried[b][a], src, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The values view of stage s at carried boundary b.
View make_view(const AdjTape* t, int s, int b) {
  View V;
  for (int a = 0; a < A_count; ++a) {
    V.val[a] = t->dummy;
    V.bar[a] = t->bar[a];
  }
  for (int a = 0; a < kNStageState; ++a) V.val[kStageState[a]] = t->stage_state[s][a];
  for (int a = 0; a < kNStageConst; ++a) V.val[kStageConst[a]] = t->stage_const[s][a];
  for (int a = 0; a < kNCarried; ++a) V.val[kCarried[a]] = t->carried[b][a];
  V.scr = t->scr;
  return V;
}

}  // namespace

hipError_t adjoint_large_step(const DGrid& g, const DNamelist& nl, const DWork& w, AdjTape* t, const DState& start,
                              const DState& bar_io, hipStream_t st) {
  if (nl.mixed) return hipErrorInvalidValue;
  int nsub[3];
  n_boundaries(nl, nsub);
  const double dt = nl.dt;
  const int rk_of[3] = {1, 2, 3};
  const double dts_of[3] = {dt / 3.0, dt / (double)nl.ns, dt / (double)nl.ns};
  const DState& S = t->scratch;
  hipError_t e;
#define ADJ_TRY(x)             \
  do {                         \
    e = (x);                   \
    if (e != hipSuccess) return e; \
  } while (0)
  // ---- forward recompute with the tape ----
  for (int a = 0; a < kNStageState; ++a)
    ADJ_TRY(copy_arr(g, kStageState[a], state_ptr(S, kStageState[a]), state_ptr(start, kStageState[a]), st));
  ADJ_TRY(copy_arr(g, A_mudf, S.mudf, start.mudf, st));
  {
    const int pairs[6][2] = {{A_u0, A_u}, {A_v0, A_v}, {A_w0, A_w}, {A_th0, A_theta}, {A_ph0, A_ph_pert}, {A_mu0, A_mu_pert}};
    for (const auto& pr : pairs) ADJ_TRY(copy_arr(g, pr[0], dwork_ptr(w, pr[0]), state_ptr(S, pr[1]), st));
  }
  std::vector<int> first_b(3);
  int b = 0;
  for (int s = 0; s < 3; ++s) {
    for (int a = 0; a < kNStageState; ++a)
      ADJ_TRY(copy_arr(g, kStageState[a], t->stage_state[s][a], state_ptr(S, kStageState[a]), st));
    ADJ_TRY(launch_stage_setup(g, nl, S, w, rk_of[s], dts_of[s], st, nullptr));
    for (int a = 0; a < kNStageConst; ++a) {
      const int arr = kStageConst[a];
      if (arr == A_phb) {
        const size_t n = ext_len(g, E_W);
        k_adj_phb<<<nblocks(n), kBs, 0, st>>>(n, S.ph_total, S.ph_pert, t->stage_const[s][a]);
        ADJ_TRY(hipGetLastError());
      } else {
        ADJ_TRY(copy_arr(g, arr, t->stage_const[s][a], dwork_ptr(w, arr), st));
      }
    }
    first_b[s] = b;
    ADJ_TRY(snap_carried(g, w, S, t, b++, true, st));
    for (int it = 0; it < nsub[s]; ++it) {
      for (int part = 0; part < 3; ++part) {
        ADJ_TRY(launch_substep_part(part, g, nl, S, w, dts_of[s], it == 0, st));
        ADJ_TRY(snap_carried(g, w, S, t, b++, it == 0 && part < 2, st));
      }
    }
    ADJ_TRY(launch_stage_finish(g, S, w, st, nullptr));
  }
  // ---- reverse ----
  ADJ_TRY(hipMemsetAsync(t->bar[0], 0,
                         (size_t)((char*)t->scr - (char*)t->bar[0]), st));  // every bar array (contiguous)
  for (int a = 0; a <= A_mudf; ++a) ADJ_TRY(copy_arr(g, a, t->bar[a], state_ptr(bar_io, a), st));
  const size_t C = (size_t)g.ny * g.nx;
  const unsigned cb = nblocks(C), fb = nblocks(n_faces(g));
  for (int s = 2; s >= 0; --s) {
    const int b0 = first_b[s];
    const int b_last = b0 + 3 * nsub[s];
    {
      View V = make_view(t, s, b_last);
      k_adj_finish_col<<<cb, kBs, 0, st>>>(g, V);
      k_adj_finish_face<<<fb, kBs, 0, st>>>(g, V);
    }
    for (int it = nsub[s] - 1; it >= 0; --it) {
      const int bs0 = b0 + 3 * it;  // before uv; +1 after uv; +2 after mu; +3 after w
      k_adj_p<<<cb, kBs, 0, st>>>(g, make_view(t, s, bs0 + 3));
      k_adj_w<<<cb, kBs, 0, st>>>(g, nl, dts_of[s], make_view(t, s, bs0 + 2));
      k_adj_mu_theta<<<cb, kBs, 0, st>>>(g, nl, dts_of[s], make_view(t, s, bs0 + 1));
      k_adj_uv<<<fb, kBs, 0, st>>>(g, dts_of[s], make_view(t, s, bs0));
    }
    {
      View V = make_view(t, s, b0);  // the arrays after the setup
      k_adj_s1<<<cb, kBs, 0, st>>>(g, nl, dts_of[s], V);
      k_adj_s2<<<cb, kBs, 0, st>>>(g, V);
      k_adj_s3<<<fb, kBs, 0, st>>>(g, V);
      k_adj_s4<<<cb, kBs, 0, st>>>(g, rk_of[s], V);
      k_adj_s4_clear<<<fb, kBs, 0, st>>>(g, V);
      k_adj_pgf<<<fb, kBs, 0, st>>>(g, V);
      k_adj_s6<<<cb, kBs, 0, st>>>(g, V);
      k_adj_adv_copy<<<nblocks((size_t)g.nz * (g.ny + g.nx)), kBs, 0, st>>>(g, V);
      for (int kind = ADV_SCALAR; kind <= ADV_W; ++kind) {
        const int nlev = kind == ADV_W ? g.nz + 1 : g.nz;
        k_adj_adv_h<<<nblocks((size_t)nlev * C), kBs, 0, st>>>(g, V, kind, nlev);
        k_adj_adv_v<<<cb, kBs, 0, st>>>(g, V, kind);
      }
      k_adj_couple_col<<<cb, kBs, 0, st>>>(g, V);
      k_adj_couple_pt<<<nblocks((size_t)g.nz * C), kBs, 0, st>>>(g, V);
    }
    ADJ_TRY(hipGetLastError());
  }
  {
    View V = make_view(t, 0, 0);
    const int pairs[6][2] = {{A_u, A_u0}, {A_v, A_v0}, {A_w, A_w0}, {A_theta, A_th0}, {A_ph_pert, A_ph0}, {A_mu_pert, A_mu0}};
    for (const auto& pr : pairs) {
      const size_t n = ext_len(g, kExtent[pr[0]]);
      k_adj_origin<<<nblocks(n), kBs, 0, st>>>(V, pr[0], pr[1], n);
    }
    ADJ_TRY(hipGetLastError());
  }
  for (int a = 0; a <= A_mudf; ++a) ADJ_TRY(copy_arr(g, a, state_ptr(bar_io, a), t->bar[a], st));
#undef ADJ_TRY
  return hipSuccess;
}

}  // namespace wd
