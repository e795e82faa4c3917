// Launch glue of the adjoint of one large step (wd_step_adjoint), shared by
// the GPU path (wrfdyn_adjoint.hip) and a host replay of the same sequence:
//  - the tape's array list and sizes (where each piece's inputs are kept),
//  - the per-piece views onto the tape (make_view),
//  - the per-thread index map of every reverse kernel (thread t of a launch
//    of nblocks(n) x kBs threads, tail threads included),
//  - the order and grid sizes of the reverse launches (reverse_sequence).
// The host replay runs reverse_sequence with a launcher that calls the
// per-thread functions for every thread of every block, so the launch glue
// is exercised on the CPU exactly as the GPU runs it.
#pragma once

#include <algorithm>
#include <vector>

#include "wrfdyn_adjoint.h"

namespace wd {
namespace adjg {

using namespace adj;

enum Ext : int { E_M, E_W, E_U, E_V, E_C, E_CU, E_CV };
// extents of the AdjArr arrays, in their order
constexpr int kExtent[A_count] = {
    E_U, E_V, E_W, E_M, E_M, E_M, E_W, E_W, E_C, E_C, E_C,
    E_U, E_V, E_W, E_M, E_W, E_C, E_M, E_M, E_W, E_M, E_U, E_V, E_W, E_M,
    E_M, E_M, E_M, E_C, E_C, E_C, E_C, E_CU, E_CU, E_CV, E_CV,
    E_U, E_V, E_W, E_M, E_W, E_M, E_M, E_M, E_W, E_M, E_M, E_M,
    E_M, E_W, E_W, E_W, E_W, E_W, E_C, E_C, E_C, E_W, E_M,
    E_M, E_C, E_W, E_W, E_W, E_W};

inline size_t ext_len(const DGrid& g, int e) {
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

// Substeps per stage (1, max(1, ns/2), ns) and the number of carried
// boundaries: one after each stage setup and one after each substep part.
inline int n_boundaries(const DNamelist& nl, int* nsub) {
  const int ns = nl.ns;
  nsub[0] = 1;
  nsub[1] = ns / 2 > 1 ? ns / 2 : 1;
  nsub[2] = ns;
  return 3 + 3 * (nsub[0] + nsub[1] + nsub[2]);
}

// Where the tape's arrays are (device memory on the GPU, host memory in the
// host replay).
struct TapePtrs {
  DState scratch{};                                   // forward recompute state
  std::vector<std::vector<double*>> stage_state;      // [3][kNStageState]
  std::vector<std::vector<double*>> stage_const;      // [3][kNStageConst]
  std::vector<std::vector<double*>> carried;          // [boundary][kNCarried]
  double* bar[A_count] = {};
  double* scr = nullptr;                              // kScrSlots planes of (nz+1, ny, nx)
  double* dummy = nullptr;                            // a zero array of the largest extent
};

struct TapeSlot {
  double** p;
  size_t n;  // doubles
};

// Every tape array, in allocation order, with the pointer it fills.  The bar
// arrays are consecutive, then scr, then dummy (the reverse clears the bars
// as one range from bar[0] to scr).
inline std::vector<TapeSlot> tape_slots(const DGrid& g, const DNamelist& nl, TapePtrs* t) {
  int nsub[3];
  const int nb = n_boundaries(nl, nsub);
  std::vector<TapeSlot> want;
  const size_t wlen = ext_len(g, E_W);
  const size_t big = std::max(ext_len(g, E_U), std::max(ext_len(g, E_V), wlen));
  double** scratch[] = {&t->scratch.u,        &t->scratch.v,        &t->scratch.w,       &t->scratch.theta,
                        &t->scratch.p_total,  &t->scratch.p_pert,   &t->scratch.ph_total, &t->scratch.ph_pert,
                        &t->scratch.mu_total, &t->scratch.mu_pert, &t->scratch.mudf};
  for (int a = 0; a <= A_mudf; ++a) want.push_back({scratch[a], ext_len(g, kExtent[a])});
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
  return want;
}

// Bytes of a slot carved from one buffer (256-byte aligned sub-buffers).
inline size_t slot_bytes(const TapeSlot& s) { return (s.n * sizeof(double) + 255) / 256 * 256; }

// The values view of stage s at carried boundary b: the stage's state and
// constants, the carried arrays of boundary b, every other array the zero
// dummy; the bars are the tape's.
inline View make_view(const TapePtrs& t, int s, int b) {
  View V;
  for (int a = 0; a < A_count; ++a) {
    V.val[a] = t.dummy;
    V.bar[a] = t.bar[a];
  }
  for (int a = 0; a < kNStageState; ++a) V.val[kStageState[a]] = t.stage_state[s][a];
  for (int a = 0; a < kNStageConst; ++a) V.val[kStageConst[a]] = t.stage_const[s][a];
  for (int a = 0; a < kNCarried; ++a) V.val[kCarried[a]] = t.carried[b][a];
  V.scr = t.scr;
  return V;
}

// ---- the reverse kernels, per thread ----
constexpr unsigned kBs = 256;
inline unsigned nblocks(size_t n) { return (unsigned)((n + kBs - 1) / kBs); }

// u faces (ny x (nx+1)) then v faces ((ny+1) x nx)
__host__ __device__ inline bool face_of(const DGrid& g, size_t f, bool* ydir, int* j, int* i) {
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
inline size_t n_faces(const DGrid& g) { return (size_t)g.ny * (g.nx + 1) + (size_t)(g.ny + 1) * g.nx; }

// The reverse kernels.
enum class AK : int {
  origin, finish_col, finish_face, p, w, mu_theta, uv, s1, s2, s3, s4, s4_clear, pgf, s6, adv_copy, adv_h, adv_v,
  couple_col, couple_pt
};

// The arguments of one reverse launch (each kernel reads its own subset).
struct RevArgs {
  DGrid g;
  DNamelist nl;
  double dts;
  View V;
  int rk_step, kind, nlev, src, dst;
  size_t n;
};

// Thread t of a launch of kernel k.
template <class Acc>
__host__ __device__ inline void rev_thread(AK k, const RevArgs& a, size_t t, Acc acc) {
  const DGrid& g = a.g;
  const size_t C = (size_t)g.ny * g.nx;
  bool yd;
  int j, i;
  switch (k) {
    case AK::origin:
      if (t < a.n) adj_origin_elem(a.V, a.src, a.dst, t);
      break;
    case AK::finish_col:
      if (t < C) adj_finish_column(g, a.V, (int)(t / g.nx), (int)(t % g.nx));
      break;
    case AK::finish_face:
      if (face_of(g, t, &yd, &j, &i)) adj_finish_face(g, a.V, yd, j, i);
      break;
    case AK::p:
      if (t < C) adj_p_column(g, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::w:
      if (t < C) adj_w_column(g, a.nl, a.dts, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::mu_theta:
      if (t < C) adj_mu_theta_column(g, a.nl, a.dts, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::uv:
      if (face_of(g, t, &yd, &j, &i)) adj_uv_face(g, a.dts, a.V, yd, j, i, acc);
      break;
    case AK::s1:
      if (t < C) adj_setup_s1_column(g, a.nl, a.dts, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::s2:
      if (t < C) adj_setup_s2_column(g, a.V, (int)(t / g.nx), (int)(t % g.nx));
      break;
    case AK::s3:
      if (face_of(g, t, &yd, &j, &i)) adj_setup_s3_face(g, a.V, yd, j, i);
      break;
    case AK::s4:
      if (t < C) adj_setup_s4_column(g, a.rk_step, a.V, (int)(t / g.nx), (int)(t % g.nx));
      break;
    case AK::s4_clear:
      if (face_of(g, t, &yd, &j, &i)) adj_setup_s4_clear(g, a.V, yd, j, i);
      break;
    case AK::pgf:
      if (face_of(g, t, &yd, &j, &i)) adj_pgf_face(g, a.V, yd, j, i, acc);
      break;
    case AK::s6:
      if (t < C) adj_setup_s6_column(g, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::adv_copy: {
      // periodic tendency copies: u rows (nz x ny) then v columns (nz x nx)
      const size_t nu = (size_t)g.nz * g.ny, nv = (size_t)g.nz * g.nx;
      if (t < nu) adj_advect_copy(g, a.V, false, (int)(t / g.ny), (int)(t % g.ny));
      else if (t < nu + nv) adj_advect_copy(g, a.V, true, (int)((t - nu) / g.nx), (int)((t - nu) % g.nx));
      break;
    }
    case AK::adv_h:
      if (t < (size_t)a.nlev * C) {
        const size_t c = t % C;
        adj_advect_h_point(g, a.V, a.kind, (int)(t / C), (int)(c / g.nx), (int)(c % g.nx), acc);
      }
      break;
    case AK::adv_v:
      if (t < C) adj_advect_v_column(g, a.V, a.kind, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::couple_col:
      if (t < C) adj_couple_column(g, a.V, (int)(t / g.nx), (int)(t % g.nx), acc);
      break;
    case AK::couple_pt:
      if (t < (size_t)g.nz * C) {
        const size_t c = t % C;
        adj_couple_point(g, a.V, (int)(t / C), (int)(c / g.nx), (int)(c % g.nx), acc);
      }
      break;
  }
}

// The reverse of one large step, stage by stage from the last: finish, the
// substeps in reverse (calc_p_rho, advance_w, advance_mu_t + theta,
// advance_uv), the stage setup's phases, then the step origin.  first_b[s]
// is the carried boundary right after stage s's setup.  launch(k, blocks,
// args) runs kernel k over blocks x kBs threads; it returns false to stop.
template <class Launch>
bool reverse_sequence(const DGrid& g, const DNamelist& nl, const TapePtrs& t, Launch& launch) {
  int nsub[3];
  n_boundaries(nl, nsub);
  const double dt = nl.dt;
  const int rk_of[3] = {1, 2, 3};
  const double dts_of[3] = {dt / 3.0, dt / (double)nl.ns, dt / (double)nl.ns};
  int first_b[3];
  first_b[0] = 0;
  for (int s = 1; s < 3; ++s) first_b[s] = first_b[s - 1] + 1 + 3 * nsub[s - 1];
  const size_t C = (size_t)g.ny * g.nx;
  const unsigned cb = nblocks(C), fb = nblocks(n_faces(g));
  RevArgs a{};
  a.g = g;
  a.nl = nl;
  for (int s = 2; s >= 0; --s) {
    const int b0 = first_b[s];
    a.dts = dts_of[s];
    a.rk_step = rk_of[s];
    a.V = make_view(t, s, b0 + 3 * nsub[s]);
    if (!launch(AK::finish_col, cb, a) || !launch(AK::finish_face, fb, a)) return false;
    for (int it = nsub[s] - 1; it >= 0; --it) {
      const int bs0 = b0 + 3 * it;  // before uv; +1 after uv; +2 after mu; +3 after w
      a.V = make_view(t, s, bs0 + 3);
      if (!launch(AK::p, cb, a)) return false;
      a.V = make_view(t, s, bs0 + 2);
      if (!launch(AK::w, cb, a)) return false;
      a.V = make_view(t, s, bs0 + 1);
      if (!launch(AK::mu_theta, cb, a)) return false;
      a.V = make_view(t, s, bs0);
      if (!launch(AK::uv, fb, a)) return false;
    }
    a.V = make_view(t, s, b0);  // the arrays after the setup
    const AK seq[] = {AK::s1, AK::s2, AK::s3, AK::s4, AK::s4_clear, AK::pgf, AK::s6};
    const unsigned grid[] = {cb, cb, fb, cb, fb, fb, cb};
    for (int q = 0; q < 7; ++q)
      if (!launch(seq[q], grid[q], a)) return false;
    if (!launch(AK::adv_copy, nblocks((size_t)g.nz * (g.ny + g.nx)), a)) return false;
    for (int kind = ADV_SCALAR; kind <= ADV_W; ++kind) {
      a.kind = kind;
      a.nlev = kind == ADV_W ? g.nz + 1 : g.nz;
      if (!launch(AK::adv_h, nblocks((size_t)a.nlev * C), a) || !launch(AK::adv_v, cb, a)) return false;
    }
    if (!launch(AK::couple_col, cb, a) || !launch(AK::couple_pt, nblocks((size_t)g.nz * C), a)) return false;
  }
  a.V = make_view(t, 0, 0);
  const int pairs[6][2] = {{A_u, A_u0}, {A_v, A_v0}, {A_w, A_w0}, {A_theta, A_th0}, {A_ph_pert, A_ph0},
                           {A_mu_pert, A_mu0}};
  for (const auto& pr : pairs) {
    a.src = pr[0];
    a.dst = pr[1];
    a.n = ext_len(g, kExtent[pr[0]]);
    if (!launch(AK::origin, nblocks(a.n), a)) return false;
  }
  return true;
}

}  // namespace adjg
}  // namespace wd
