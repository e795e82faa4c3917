// Serial host run of the adjoint piece bodies of wrf_gpu_amd/csrc/wrfdyn_adjoint.h,
// for checking them on the CPU against the gradient oracle's per-piece VJPs
// (tests/test_adjoint_pieces.py).  Test infrastructure: the product runs the
// same bodies only in the GPU kernels of wrfdyn_adjoint.hip.
#include "wrfdyn_adjoint.h"

namespace {

// Layouts of orc_grid / orc_namelist (oracle/wrfdyn_oracle.h), which the
// Python oracle binding fills.
struct HostGrid {
  int nx, ny, nz;
  double dx, dy;
  const double *c1h, *c2h, *c1f, *c2f, *dnw, *rdnw, *dn, *rdn, *fnm, *fnp;
  double cf1, cf2, cf3;
  const double *msftx, *msfty, *msfux, *msfuy, *msfvx, *msfvy;
};
struct HostNamelist {
  double dt;
  int ns;
  double epssm;
  int top_lid, w_damping, damp_opt;
  double dampcoef, zdamp;
  int diff_6th_opt;
  double diff_6th_factor;
};

struct PlainAdd {
  void operator()(double* p, double v) const { *p += v; }
};

wd::DGrid to_dgrid(const HostGrid& h) {
  wd::DGrid g{};
  g.nx = h.nx; g.ny = h.ny; g.nz = h.nz;
  g.dx = h.dx; g.dy = h.dy; g.rdx = 1.0 / h.dx; g.rdy = 1.0 / h.dy;
  g.c1h = h.c1h; g.c2h = h.c2h; g.c1f = h.c1f; g.c2f = h.c2f;
  g.dnw = h.dnw; g.rdnw = h.rdnw; g.dn = h.dn; g.rdn = h.rdn; g.fnm = h.fnm; g.fnp = h.fnp;
  g.cf1 = h.cf1; g.cf2 = h.cf2; g.cf3 = h.cf3;
  const double dn_top = h.dn[h.nz - 1];
  const double dn_safe = fabs(dn_top) > 1.0e-30 ? dn_top : 1.0;
  g.cfn = (0.5 * h.dnw[h.nz - 1] + h.dn[h.nz - 1]) / dn_safe;
  g.cfn1 = -0.5 * h.dnw[h.nz - 1] / dn_safe;
  g.msftx = h.msftx; g.msfty = h.msfty; g.msfux = h.msfux; g.msfuy = h.msfuy;
  g.msfvx = h.msfvx; g.msfvy = h.msfvy;
  g.xcd_remap = 0;
  return g;
}

wd::DNamelist to_dnl(const HostNamelist& h) {
  wd::DNamelist n{};
  n.dt = h.dt; n.ns = h.ns; n.epssm = h.epssm; n.top_lid = h.top_lid; n.w_damping = h.w_damping;
  n.damp_opt = h.damp_opt; n.dampcoef = h.dampcoef; n.zdamp = h.zdamp; n.mixed = 0;
  return n;
}

// The stage setup's reverse phases S1 .. S9, each over the whole grid.
int setup(const wd::DGrid& g, const wd::DNamelist& nl, int rk_step, double dts, const wd::adj::View& V, PlainAdd acc) {
  using namespace wd::adj;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s1_column(g, nl, dts, V, j, i, acc);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s2_column(g, V, j, i);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) adj_setup_s3_face(g, V, false, j, i);
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s3_face(g, V, true, j, i);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s4_column(g, rk_step, V, j, i);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) adj_setup_s4_clear(g, V, false, j, i);
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s4_clear(g, V, true, j, i);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) adj_pgf_face(g, V, false, j, i, acc);
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) adj_pgf_face(g, V, true, j, i, acc);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) adj_setup_s6_column(g, V, j, i, acc);
  for (int k = 0; k < nz; ++k) {
    for (int r = 0; r < ny; ++r) adj_advect_copy(g, V, false, k, r);
    for (int r = 0; r < nx; ++r) adj_advect_copy(g, V, true, k, r);
  }
  for (int kind = ADV_SCALAR; kind <= ADV_W; ++kind) {
    const int kmax = kind == ADV_W ? nz : nz - 1;
    for (int k = 0; k <= kmax; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) adj_advect_h_point(g, V, kind, k, j, i, acc);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) adj_advect_v_column(g, V, kind, j, i, acc);
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) adj_couple_column(g, V, j, i, acc);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) adj_couple_point(g, V, k, j, i, acc);
  return 0;
}

}  // namespace

extern "C" {

int adjh_count(void) { return wd::adj::A_count; }

// Adjoint of one substep piece (2 uv, 3 mu_theta, 4 w, 5 p; the numbering of
// the oracle's ORC_PIECE_*).  Returns 0, or -2 for a piece not built here.
int adjh_piece(const HostGrid* hg, const HostNamelist* hn, int piece, int rk_step, double dts,
               const double* const* val, double* const* bar, double* scr) {
  const wd::DGrid g = to_dgrid(*hg);
  const wd::DNamelist nl = to_dnl(*hn);
  wd::adj::View V;
  for (int a = 0; a < wd::adj::A_count; ++a) {
    V.val[a] = val[a];
    V.bar[a] = bar[a];
  }
  V.scr = scr;
  PlainAdd acc;
  const size_t nx = g.nx, ny = g.ny, nz = g.nz;
  switch (piece) {
    case 0: {
      const int pairs[6][2] = {{wd::adj::A_u, wd::adj::A_u0},         {wd::adj::A_v, wd::adj::A_v0},
                               {wd::adj::A_w, wd::adj::A_w0},         {wd::adj::A_theta, wd::adj::A_th0},
                               {wd::adj::A_ph_pert, wd::adj::A_ph0},  {wd::adj::A_mu_pert, wd::adj::A_mu0}};
      const size_t len[6] = {nz * ny * (nx + 1), nz * (ny + 1) * nx, (nz + 1) * ny * nx, nz * ny * nx,
                             (nz + 1) * ny * nx, ny * nx};
      for (int p = 0; p < 6; ++p)
        for (size_t q = 0; q < len[p]; ++q) wd::adj::adj_origin_elem(V, pairs[p][0], pairs[p][1], q);
      return 0;
    }
    case 1:
      return setup(g, nl, rk_step, dts, V, acc);
    case 6:
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_finish_column(g, V, j, i);
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i <= g.nx; ++i) wd::adj::adj_finish_face(g, V, false, j, i);
      for (int j = 0; j <= g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_finish_face(g, V, true, j, i);
      return 0;
    case 2:
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i <= g.nx; ++i) wd::adj::adj_uv_face(g, dts, V, false, j, i, acc);
      for (int j = 0; j <= g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_uv_face(g, dts, V, true, j, i, acc);
      return 0;
    case 3:
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_mu_theta_column(g, nl, dts, V, j, i, acc);
      return 0;
    case 4:
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_w_column(g, nl, dts, V, j, i, acc);
      return 0;
    case 5:
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) wd::adj::adj_p_column(g, V, j, i, acc);
      return 0;
    default:
      return -2;
  }
}

}  // extern "C"
