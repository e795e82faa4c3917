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

#include "wrfdyn_adjoint_glue.h"
#include "wrfdyn_kernels.h"

namespace wd {

using namespace adj;
using namespace adjg;

namespace {

struct AtomicAdd {
  __device__ void operator()(double* p, double v) const { atomicAdd(p, v); }
};

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

__global__ void __launch_bounds__(kBs) k_adj_phb(size_t n, const double* pht, const double* php, double* out) {
  const size_t q = (size_t)blockIdx.x * kBs + threadIdx.x;
  if (q < n) out[q] = pht[q] - php[q];
}

// One reverse kernel: thread t of the launch runs rev_thread (the per-thread
// index map shared with the host replay), bars shared by neighbouring
// threads through atomic adds.
template <AK K>
__global__ void __launch_bounds__(kBs) k_adj_rev(RevArgs a) {
  rev_thread(K, a, (size_t)blockIdx.x * kBs + threadIdx.x, AtomicAdd{});
}

struct GpuLaunch {
  hipStream_t st;
  hipError_t err = hipSuccess;
  bool operator()(AK k, unsigned nb, const RevArgs& a) {
    switch (k) {
      case AK::origin: k_adj_rev<AK::origin><<<nb, kBs, 0, st>>>(a); break;
      case AK::finish_col: k_adj_rev<AK::finish_col><<<nb, kBs, 0, st>>>(a); break;
      case AK::finish_face: k_adj_rev<AK::finish_face><<<nb, kBs, 0, st>>>(a); break;
      case AK::p: k_adj_rev<AK::p><<<nb, kBs, 0, st>>>(a); break;
      case AK::w: k_adj_rev<AK::w><<<nb, kBs, 0, st>>>(a); break;
      case AK::mu_theta: k_adj_rev<AK::mu_theta><<<nb, kBs, 0, st>>>(a); break;
      case AK::uv: k_adj_rev<AK::uv><<<nb, kBs, 0, st>>>(a); break;
      case AK::s1: k_adj_rev<AK::s1><<<nb, kBs, 0, st>>>(a); break;
      case AK::s2: k_adj_rev<AK::s2><<<nb, kBs, 0, st>>>(a); break;
      case AK::s3: k_adj_rev<AK::s3><<<nb, kBs, 0, st>>>(a); break;
      case AK::s4: k_adj_rev<AK::s4><<<nb, kBs, 0, st>>>(a); break;
      case AK::s4_clear: k_adj_rev<AK::s4_clear><<<nb, kBs, 0, st>>>(a); break;
      case AK::pgf: k_adj_rev<AK::pgf><<<nb, kBs, 0, st>>>(a); break;
      case AK::s6: k_adj_rev<AK::s6><<<nb, kBs, 0, st>>>(a); break;
      case AK::adv_copy: k_adj_rev<AK::adv_copy><<<nb, kBs, 0, st>>>(a); break;
      case AK::adv_h: k_adj_rev<AK::adv_h><<<nb, kBs, 0, st>>>(a); break;
      case AK::adv_v: k_adj_rev<AK::adv_v><<<nb, kBs, 0, st>>>(a); break;
      case AK::couple_col: k_adj_rev<AK::couple_col><<<nb, kBs, 0, st>>>(a); break;
      case AK::couple_pt: k_adj_rev<AK::couple_pt><<<nb, kBs, 0, st>>>(a); break;
    }
    err = hipGetLastError();
    return err == hipSuccess;
  }
};

}  // namespace

struct AdjTape {
  void* buf = nullptr;
  size_t bytes = 0;
  TapePtrs p;  // every array carved from buf
};

hipError_t adjoint_tape_create(const DGrid& g, const DNamelist& nl, AdjTape** out, size_t* bytes) {
  *out = nullptr;
  AdjTape* t = new (std::nothrow) AdjTape();
  if (!t) return hipErrorOutOfMemory;
  const std::vector<TapeSlot> want = tape_slots(g, nl, &t->p);
  size_t total = 0;
  for (const TapeSlot& x : want) total += slot_bytes(x);
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
  for (const TapeSlot& x : want) {
    *x.p = reinterpret_cast<double*>(base + off);
    off += slot_bytes(x);
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
hipError_t snap_carried(const DGrid& g, const DWork& w, const DState& s, TapePtrs& t, int b, bool first,
                        hipStream_t st) {
  for (int a = 0; a < kNCarried; ++a) {
    const int arr = kCarried[a];
    const double* src = arr == A_mudf ? s.mudf : dwork_ptr(w, arr);
    hipError_t e;
    if (first && arr == A_t2ave)
      e = hipMemsetAsync(t.carried[b][a], 0, ext_len(g, E_M) * sizeof(double), st);
    else if (first && arr == A_pm1)
      e = copy_arr(g, arr, t.carried[b][a], w.p, st);
    else
      e = copy_arr(g, arr, t.carried[b][a], src, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t adjoint_large_step(const DGrid& g, const DNamelist& nl, const DWork& w, AdjTape* tape, const DState& start,
                              const DState& bar_io, hipStream_t st) {
  if (nl.mixed) return hipErrorInvalidValue;
  TapePtrs& t = tape->p;
  int nsub[3];
  n_boundaries(nl, nsub);
  const double dt = nl.dt;
  const int rk_of[3] = {1, 2, 3};
  const double dts_of[3] = {dt / 3.0, dt / (double)nl.ns, dt / (double)nl.ns};
  const DState& S = t.scratch;
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
  int b = 0;
  for (int s = 0; s < 3; ++s) {
    for (int a = 0; a < kNStageState; ++a)
      ADJ_TRY(copy_arr(g, kStageState[a], t.stage_state[s][a], state_ptr(S, kStageState[a]), st));
    ADJ_TRY(launch_stage_setup(g, nl, S, w, rk_of[s], dts_of[s], st, nullptr));
    for (int a = 0; a < kNStageConst; ++a) {
      const int arr = kStageConst[a];
      if (arr == A_phb) {
        const size_t n = ext_len(g, E_W);
        k_adj_phb<<<nblocks(n), kBs, 0, st>>>(n, S.ph_total, S.ph_pert, t.stage_const[s][a]);
        ADJ_TRY(hipGetLastError());
      } else {
        ADJ_TRY(copy_arr(g, arr, t.stage_const[s][a], dwork_ptr(w, arr), st));
      }
    }
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
  ADJ_TRY(hipMemsetAsync(t.bar[0], 0, (size_t)((char*)t.scr - (char*)t.bar[0]), st));  // every bar array
  for (int a = 0; a <= A_mudf; ++a) ADJ_TRY(copy_arr(g, a, t.bar[a], state_ptr(bar_io, a), st));
  GpuLaunch L{st};
  if (!reverse_sequence(g, nl, t, L)) return L.err;
  for (int a = 0; a <= A_mudf; ++a) ADJ_TRY(copy_arr(g, a, state_ptr(bar_io, a), t.bar[a], st));
#undef ADJ_TRY
  return hipSuccess;
}

}  // namespace wd
