// Adjoint (reverse-mode) bodies of the acoustic substep pieces of the large
// step, one column (or one u/v face column) per call.
//
// Each function takes the values of every array the piece reads, as they
// were BEFORE the piece ran, and the adjoints (bars) of the arrays AFTER it,
// and turns the bars into those before it: the bar of an array the piece
// overwrites becomes the part from the piece's own reads of its old value,
// and the bars of the arrays it reads gain their contributions.  That is the
// vector-Jacobian product of the piece, taken statement by statement in
// reverse with the same derivative conventions as jax: sign() has derivative
// 0, |x| has sign(x), and a select passes the chosen operand's derivative.
//
// The pieces and the reference lines they differentiate:
//   adj_p_column      calc_p_rho_step, core/calc_p_rho.py:142-183
//   adj_w_column      advance_w_wrf, core/advance_w.py:293-627 (Thomas,
//                     damp_opt=3, w_damping, the terrain surface w)
//   adj_mu_theta_column  advance_mu_t_wrf periodic with the theta update,
//                     dynamics/mu_t_advance.py:139-292
//   adj_uv_u_face / adj_uv_v_face  advance_uv_wrf, core/acoustic.py:626-786
//
// Arrays are named by AdjArr, the state fields and the step's work arrays.
// Bars of arrays that several columns touch are updated through the Acc
// functor (an atomic add in the kernels, a plain add in a serial host run);
// every update of such an array goes through it.  The functions are
// __host__ __device__ so the same source runs in the GPU kernels
// (wrfdyn_adjoint.hip) and in a serial host check against the oracle.
#pragma once

#include "wrfdyn_common.h"

namespace wd {
namespace adj {

enum AdjArr : int {
  A_u, A_v, A_w, A_theta, A_p_total, A_p_pert, A_ph_total, A_ph_pert, A_mu_total, A_mu_pert, A_mudf,
  A_u0, A_v0, A_w0, A_th0, A_ph0, A_mu0, A_ru, A_rv, A_rom, A_divv, A_ut, A_vt, A_wt, A_tht,
  A_alt_raw, A_al_ls, A_php_ls, A_mub, A_mut, A_muts0, A_mu_work0, A_muu, A_muus, A_muv, A_muvs,
  A_u_work, A_v_work, A_w_work, A_th_work, A_ph_work, A_c2a, A_alt, A_pb, A_phb, A_php, A_p, A_al,
  A_pm1, A_rw_tend, A_ph_tend, A_ca, A_calpha, A_cgamma, A_mu, A_muts, A_muave, A_ww, A_t2ave,
  A_dvdxi, A_mudf_new, A_ww_new, A_wsolve, A_rhs, A_tmp_ph_total,
  A_count
};

// Values before the piece (val), bars (bar, updated in place), and column
// scratch for recomputed intermediates and their bars (scr: kScrSlots
// planes of (nz+1, ny, nx) doubles).
struct View {
  const double* val[A_count];
  double* bar[A_count];
  double* scr;
};
constexpr int kScrSlots = 10;

__host__ __device__ inline double sgn0(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }
__host__ __device__ inline int wrap_h(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}
__host__ __device__ inline int clamp_h(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// ---------------------------------------------------------------------------
// calc_p_rho_step (smdiv) on column (j, i).
// ---------------------------------------------------------------------------
template <class Acc>
__host__ __device__ inline void adj_p_column(const DGrid& g, const View& V, int j, int i, Acc acc) {
  (void)acc;
  const int nz = g.nz;
  const size_t c = i2(g, j, i);
  const double muts = V.val[A_muts][c], mut = V.val[A_mut][c];
  double b_muts = 0.0, b_mut = 0.0, b_phw_up = 0.0;  // b_phw_up: bar of ph_work[k+1] from level k
  for (int k = nz - 1; k >= 0; --k) {
    const size_t q = im(g, k, j, i), q1 = im(g, k + 1, j, i);
    const double alt = V.val[A_alt][q], c2a = V.val[A_c2a][q], thw = V.val[A_th_work][q];
    const double phk1 = V.val[A_ph_work][q1], phk = V.val[A_ph_work][q];
    const double theta = V.val[A_theta][q];
    // forward
    const double mu_work = muts - mut;
    const double mass_h = g.c1h[k] * muts + g.c2h[k];
    const bool mass_ok = fabs(mass_h) > 1.0e-12;
    const double safe_mass = mass_ok ? mass_h : 1.0e-12;
    const double mu_term = g.c1h[k] * mu_work;
    const double A = alt * mu_term + g.rdnw[k] * (phk1 - phk);
    const double al = -A / safe_mass;
    const double theta_1 = theta - kT0;
    const double ttr = kT0 + theta_1;
    const bool t_ok = fabs(ttr) > 1.0e-6;
    const double safe_t = t_ok ? ttr : 1.0e-6;
    const double Nm = alt * (thw - mu_term * theta_1);
    const double D = safe_mass * safe_t;
    const double X = Nm / D;
    // reverse: al, p, pm1 are overwritten
    const double b_al = V.bar[A_al][q], b_p = V.bar[A_p][q], b_pm1 = V.bar[A_pm1][q];
    V.bar[A_al][q] = 0.0;
    V.bar[A_p][q] = 0.0;
    V.bar[A_pm1][q] = -kSmdiv * b_p;  // p = p0 + smdiv (p0 - pm1_old)
    const double b_p0 = b_p * (1.0 + kSmdiv) + b_pm1;
    double b_alt = 0.0, b_mu_term = 0.0, b_safe_mass = 0.0, b_theta_1 = 0.0, b_thw = 0.0;
    // p0 = c2a (X - al)
    V.bar[A_c2a][q] += b_p0 * (X - al);
    const double b_X = b_p0 * c2a;
    const double b_al_t = b_al - b_p0 * c2a;
    const double b_N = b_X / D, b_D = -b_X * X / D;
    b_alt += b_N * (thw - mu_term * theta_1);
    b_thw += b_N * alt;
    b_mu_term -= b_N * alt * theta_1;
    b_theta_1 -= b_N * alt * mu_term;
    b_safe_mass += b_D * safe_t;
    const double b_safe_t = b_D * safe_mass;
    if (t_ok) b_theta_1 += b_safe_t;
    // al = -A / safe_mass
    const double b_A = -b_al_t / safe_mass;
    b_safe_mass += b_al_t * A / (safe_mass * safe_mass);
    b_alt += b_A * mu_term;
    b_mu_term += b_A * alt;
    b_phw_up += b_A * g.rdnw[k];  // ph_work[k+1]
    const double b_phk = -b_A * g.rdnw[k];
    V.bar[A_ph_work][q1] += b_phw_up;
    b_phw_up = b_phk;  // carried to level k (added when k is done)
    if (mass_ok) b_muts += b_safe_mass * g.c1h[k];
    const double b_mu_work = b_mu_term * g.c1h[k];
    b_muts += b_mu_work;
    b_mut -= b_mu_work;
    V.bar[A_alt][q] += b_alt;
    V.bar[A_th_work][q] += b_thw;
    V.bar[A_theta][q] += b_theta_1;
  }
  V.bar[A_ph_work][im(g, 0, j, i)] += b_phw_up;
  V.bar[A_muts][c] += b_muts;
  V.bar[A_mut][c] += b_mut;
}

// ---------------------------------------------------------------------------
// advance_w on column (j, i): t_2ave, the phi-equation rhs, the explicit w
// with w_damping, the acoustic terms, the terrain surface w, the Thomas
// sweeps, Rayleigh damping (damp_opt=3) and the geopotential finish.
// Scratch planes: 0 t2n, 1 rhs before the mass scaling, 2 rhs after it,
// 3 w before Thomas, 4 the forward sweep, 5 the back sweep, 6 w after
// damping; 7, 8, 9 the bars of the w chain, of rhs and of t2n.
// ---------------------------------------------------------------------------
template <class Acc>
__host__ __device__ inline void adj_w_column(const DGrid& g, const DNamelist& nl, double dts, const View& V, int j,
                                             int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const size_t c = i2(g, j, i), plane = (size_t)ny * nx;
  double* const scr = V.scr;
  auto S = [=](int s, int f) -> double& { return scr[((size_t)s * (nz + 1) + f) * plane + c]; };
  auto Q = [=](int f) { return im(g, f, j, i); };
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double msfty = g.msfty[c], msftx = g.msftx[c];
  const double msft_inv = 1.0 / msfty;
  const double mut = val[A_mut][c], muts = val[A_muts][c], muave = val[A_muave][c];
  auto WW = [=](int f) { return val[A_ww][Q(f)]; };
  auto WQ = [=](int f) { return val[A_w_work][Q(f)]; };
  auto PHW = [=](int f) { return val[A_ph_work][Q(f)]; };
  auto PH1T = [=](int f) { return val[A_ph0][Q(f)] + val[A_phb][Q(f)]; };
  auto MF = [=](int f) { return g.c1f[f] * mut + g.c2f[f]; };
  const double c05geps = (0.5 * kG) * eps_m;
  const double c05dtg = (0.5 * dts) * kG;

  // ---- forward recompute ----
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k);
    const double mass_h = g.c1h[k] * muts + g.c2h[k];
    const double half = 0.5 * (eps_p * val[A_th_work][q] + eps_m * val[A_t2ave][q]);
    const double ttr = kT0 + (val[A_theta][q] - kT0);
    S(0, k) = (half + (g.c1h[k] * muave) * kT0) / (mass_h * ttr);
  }
  for (int f = 1; f <= nz; ++f) {
    double r = dts * (val[A_ph_tend][Q(f)] + c05geps * WQ(f));
    if (f < nz) {
      const double wd_f1 = ((0.5 * (WW(f + 1) + WW(f))) * g.rdnw[f]) * (PH1T(f + 1) - PH1T(f));
      const double wd_f = ((0.5 * (WW(f) + WW(f - 1))) * g.rdnw[f - 1]) * (PH1T(f) - PH1T(f - 1));
      const double adv = dts * (g.fnm[f] * wd_f1 + g.fnp[f] * wd_f);
      r = r + (-adv);
    }
    S(1, f) = r;
  }
  S(2, 0) = 0.0;
  for (int f = 1; f <= nz; ++f) S(2, f) = PHW(f) + (msfty * S(1, f)) / MF(f);
  if (nl.top_lid) S(2, nz) = 0.0;
  // terrain surface w
  const int jn = clamp_h(j + 1, 0, ny - 1), js = clamp_h(j - 1, 0, ny - 1);
  const int ie = clamp_h(i + 1, 0, nx - 1), iw = clamp_h(i - 1, 0, nx - 1);
  const double* phb = val[A_phb];
  const double ht = phb[i2(g, j, i)] / kG;
  const double ht_n = phb[i2(g, jn, i)] / kG, ht_s = phb[i2(g, js, i)] / kG;
  const double ht_e = phb[i2(g, j, ie)] / kG, ht_w = phb[i2(g, j, iw)] / kG;
  const double dy_n = ht_n - ht, dy_s = ht - ht_s, dx_e = ht_e - ht, dx_w = ht - ht_w;
  const double* u = val[A_u];
  const double* v = val[A_v];
  const double v_cf_n = (g.cf1 * v[iv(g, 0, j + 1, i)] + g.cf2 * v[iv(g, 1, j + 1, i)]) + g.cf3 * v[iv(g, 2, j + 1, i)];
  const double v_cf_s = (g.cf1 * v[iv(g, 0, j, i)] + g.cf2 * v[iv(g, 1, j, i)]) + g.cf3 * v[iv(g, 2, j, i)];
  const double u_cf_e = (g.cf1 * u[iu(g, 0, j, i + 1)] + g.cf2 * u[iu(g, 1, j, i + 1)]) + g.cf3 * u[iu(g, 2, j, i + 1)];
  const double u_cf_w = (g.cf1 * u[iu(g, 0, j, i)] + g.cf2 * u[iu(g, 1, j, i)]) + g.cf3 * u[iu(g, 2, j, i)];
  const double Ky = (msfty * 0.5) * g.rdy, Kx = (msftx * 0.5) * g.rdx;
  S(3, 0) = Ky * (dy_n * v_cf_n + dy_s * v_cf_s) + Kx * (dx_e * u_cf_e + dx_w * u_cf_w);
  S(3, nz) = 0.0;
  for (int f = 1; f < nz; ++f) {
    double rw = val[A_rw_tend][Q(f)];
    if (nl.w_damping == 1) {
      const double mass_f = MF(f);
      const double safe_mf = fabs(mass_f) > 1.0e-12 ? mass_f : 1.0e-12;
      const double vert_cfl = fabs(((WW(f) / safe_mf) * g.rdnw[f]) * dts);
      const double damp = (((-sgn0(WQ(f))) * kWAlpha) * (vert_cfl - kWBeta)) * mass_f;
      rw = rw + (vert_cfl > kWBeta ? damp : 0.0);
    }
    double wn = WQ(f) + dts * rw;
    const double mh_u0 = g.c1h[f] * mut + g.c2h[f], mh_l0 = g.c1h[f - 1] * mut + g.c2h[f - 1];
    const double mh_u = fabs(mh_u0) > 1.0e-12 ? mh_u0 : 1.0e-12;
    const double mh_l = fabs(mh_l0) > 1.0e-12 ? mh_l0 : 1.0e-12;
    const double coef_u = (val[A_c2a][Q(f)] * g.rdnw[f]) / mh_u;
    const double coef_l = (val[A_c2a][Q(f - 1)] * g.rdnw[f - 1]) / mh_l;
    const double tA_u = coef_u * (eps_p * (S(2, f + 1) - S(2, f)) + eps_m * (PHW(f + 1) - PHW(f)));
    const double tA_l = coef_l * (eps_p * (S(2, f) - S(2, f - 1)) + eps_m * (PHW(f) - PHW(f - 1)));
    const double termA = ((msft_inv * 1.0) * (c05dtg * g.rdn[f])) * (tA_u - tA_l);
    const double bu = (val[A_c2a][Q(f)] * val[A_alt][Q(f)]) * S(0, f);
    const double bl = (val[A_c2a][Q(f - 1)] * val[A_alt][Q(f - 1)]) * S(0, f - 1);
    const double termB = ((dts * kG) * msft_inv) * (g.rdn[f] * (bu - bl) - (g.c1f[f] * muave));
    S(3, f) = wn + (termA + termB);
  }
  S(4, 0) = S(3, 0);
  for (int f = 1; f <= nz; ++f) S(4, f) = (S(3, f) - val[A_ca][Q(f)] * S(4, f - 1)) * val[A_calpha][Q(f)];
  S(5, nz) = S(4, nz);
  for (int f = nz - 1; f >= 1; --f) S(5, f) = S(4, f) - val[A_cgamma][Q(f)] * S(5, f + 1);
  S(5, 0) = S(4, 0);
  const bool damping = nl.damp_opt == 3 && nl.dampcoef > 0.0;
  const double dampmag = dts * nl.dampcoef, hdepth = nl.zdamp;
  const double htop = PH1T(nz) / kG;
  const double hbot = htop - hdepth;
  for (int f = 0; f <= nz; ++f) S(6, f) = S(5, f);
  if (damping) {
    for (int f = 1; f <= nz; ++f) {
      const double hk = PH1T(f) / kG;
      const double sn = wd_sin(((0.5 * kPi) * (hk - hbot)) / hdepth);
      const double dampwt = hk >= hbot ? dampmag * (sn * sn) : 0.0;
      S(6, f) = (S(5, f) - (dampwt * MF(f)) * val[A_w][Q(f)]) / (1.0 + dampwt);
    }
  }

  // ---- reverse ----
  double b_mut = 0.0, b_muts = 0.0, b_muave = 0.0;
  // outputs: w_work (every face), ph_work (faces 1..nz), t2ave (every level)
  for (int f = 0; f <= nz; ++f) {
    S(7, f) = bar[A_w_work][Q(f)];
    bar[A_w_work][Q(f)] = 0.0;
    S(8, f) = 0.0;
  }
  for (int k = 0; k < nz; ++k) {
    S(9, k) = bar[A_t2ave][Q(k)];
    bar[A_t2ave][Q(k)] = 0.0;
  }
  const double Kph = (((msfty * 0.5) * dts) * kG) * eps_p;
  for (int f = 1; f <= nz; ++f) {
    const double b = bar[A_ph_work][Q(f)];
    bar[A_ph_work][Q(f)] = 0.0;
    const double mfs = g.c1f[f] * muts + g.c2f[f];
    S(8, f) += b;
    S(7, f) += b * Kph / mfs;
    b_muts += (-b * (Kph * S(6, f)) / (mfs * mfs)) * g.c1f[f];
  }
  // Rayleigh damping: S(7) becomes the bar of the back sweep
  if (damping) {
    double b_htop = 0.0;
    for (int f = 1; f <= nz; ++f) {
      const double hk = PH1T(f) / kG;
      if (!(hk >= hbot)) continue;  // dampwt = 0: w passes unchanged
      const double arg = ((0.5 * kPi) * (hk - hbot)) / hdepth;
      const double sn = wd_sin(arg);
      const double dampwt = dampmag * (sn * sn);
      const double mass_f = MF(f);
      const double wsv = val[A_w][Q(f)];
      const double den = 1.0 + dampwt;
      const double b_wd = S(7, f);
      const double b_num = b_wd / den;
      const double b_den = -b_wd * S(6, f) / den;
      const double b_dm = -b_num * wsv;
      acc(&bar[A_w][Q(f)], -b_num * (dampwt * mass_f));
      const double b_dampwt = b_dm * mass_f + b_den;
      b_mut += (b_dm * dampwt) * g.c1f[f];
      const double b_arg = (b_dampwt * dampmag) * (2.0 * sn) * cos(arg);
      const double b_hk = b_arg * (0.5 * kPi) / hdepth;
      const double b_ph1t = b_hk / kG;
      bar[A_ph0][Q(f)] += b_ph1t;
      acc(&bar[A_phb][Q(f)], b_ph1t);
      b_htop -= b_hk;  // hbot = htop - hdepth
      S(7, f) = b_num;
    }
    const double b_ph1t_top = b_htop / kG;
    bar[A_ph0][Q(nz)] += b_ph1t_top;
    acc(&bar[A_phb][Q(nz)], b_ph1t_top);
  }
  // back sweep x[f] = y[f] - cgamma[f] x[f+1]: S(7) becomes the bar of y
  for (int f = 1; f <= nz - 1; ++f) {
    const double bx = S(7, f);
    bar[A_cgamma][Q(f)] += -bx * S(5, f + 1);
    S(7, f + 1) += -bx * val[A_cgamma][Q(f)];
  }
  // forward sweep y[f] = (wE[f] - ca[f] y[f-1]) alpha[f]: S(7) becomes the bar of wE
  for (int f = nz; f >= 1; --f) {
    const double by = S(7, f);
    const double t = S(3, f) - val[A_ca][Q(f)] * S(4, f - 1);
    const double bt = by * val[A_calpha][Q(f)];
    bar[A_calpha][Q(f)] += by * t;
    S(7, f) = bt;
    bar[A_ca][Q(f)] += -bt * S(4, f - 1);
    S(7, f - 1) += -bt * val[A_ca][Q(f)];
  }
  // surface w
  {
    const double b_ws = S(7, 0);
    const double b_sy = b_ws * Ky, b_sx = b_ws * Kx;
    const double b_dy_n = b_sy * v_cf_n, b_dy_s = b_sy * v_cf_s;
    const double b_dx_e = b_sx * u_cf_e, b_dx_w = b_sx * u_cf_w;
    const double b_vn = b_sy * dy_n, b_vs = b_sy * dy_s, b_ue = b_sx * dx_e, b_uw = b_sx * dx_w;
    const double b_ht = (-b_dy_n + b_dy_s) + (-b_dx_e + b_dx_w);
    acc(&bar[A_phb][i2(g, j, i)], b_ht / kG);
    acc(&bar[A_phb][i2(g, jn, i)], b_dy_n / kG);
    acc(&bar[A_phb][i2(g, js, i)], -b_dy_s / kG);
    acc(&bar[A_phb][i2(g, j, ie)], b_dx_e / kG);
    acc(&bar[A_phb][i2(g, j, iw)], -b_dx_w / kG);
    const double cf[3] = {g.cf1, g.cf2, g.cf3};
    for (int k = 0; k < 3; ++k) {
      acc(&bar[A_v][iv(g, k, j + 1, i)], b_vn * cf[k]);
      acc(&bar[A_v][iv(g, k, j, i)], b_vs * cf[k]);
      acc(&bar[A_u][iu(g, k, j, i + 1)], b_ue * cf[k]);
      acc(&bar[A_u][iu(g, k, j, i)], b_uw * cf[k]);
    }
  }
  // interior faces: wE = w_work + dts rw + termA + termB
  for (int f = 1; f < nz; ++f) {
    const double bw = S(7, f);
    // termB
    const double b_inner = bw * ((dts * kG) * msft_inv);
    const double b_bu = b_inner * g.rdn[f], b_bl = -b_inner * g.rdn[f];
    b_muave += -b_inner * g.c1f[f];
    const double c2a_u = val[A_c2a][Q(f)], c2a_l = val[A_c2a][Q(f - 1)];
    const double alt_u = val[A_alt][Q(f)], alt_l = val[A_alt][Q(f - 1)];
    bar[A_c2a][Q(f)] += b_bu * alt_u * S(0, f);
    bar[A_alt][Q(f)] += b_bu * c2a_u * S(0, f);
    S(9, f) += b_bu * (c2a_u * alt_u);
    bar[A_c2a][Q(f - 1)] += b_bl * alt_l * S(0, f - 1);
    bar[A_alt][Q(f - 1)] += b_bl * c2a_l * S(0, f - 1);
    S(9, f - 1) += b_bl * (c2a_l * alt_l);
    // termA
    const double CA = (msft_inv * 1.0) * (c05dtg * g.rdn[f]);
    const double b_tAu = bw * CA, b_tAl = -bw * CA;
    const double mh_u0 = g.c1h[f] * mut + g.c2h[f], mh_l0 = g.c1h[f - 1] * mut + g.c2h[f - 1];
    const bool u_ok = fabs(mh_u0) > 1.0e-12, l_ok = fabs(mh_l0) > 1.0e-12;
    const double mh_u = u_ok ? mh_u0 : 1.0e-12, mh_l = l_ok ? mh_l0 : 1.0e-12;
    const double coef_u = (c2a_u * g.rdnw[f]) / mh_u, coef_l = (c2a_l * g.rdnw[f - 1]) / mh_l;
    const double Eu = eps_p * (S(2, f + 1) - S(2, f)) + eps_m * (PHW(f + 1) - PHW(f));
    const double El = eps_p * (S(2, f) - S(2, f - 1)) + eps_m * (PHW(f) - PHW(f - 1));
    const double b_cu = b_tAu * Eu, b_cl = b_tAl * El;
    const double b_Eu = b_tAu * coef_u, b_El = b_tAl * coef_l;
    S(8, f + 1) += b_Eu * eps_p;
    S(8, f) += -b_Eu * eps_p + b_El * eps_p;
    S(8, f - 1) += -b_El * eps_p;
    bar[A_ph_work][Q(f + 1)] += b_Eu * eps_m;
    bar[A_ph_work][Q(f)] += -b_Eu * eps_m + b_El * eps_m;
    bar[A_ph_work][Q(f - 1)] += -b_El * eps_m;
    bar[A_c2a][Q(f)] += b_cu * g.rdnw[f] / mh_u;
    bar[A_c2a][Q(f - 1)] += b_cl * g.rdnw[f - 1] / mh_l;
    if (u_ok) b_mut += (-b_cu * coef_u / mh_u) * g.c1h[f];
    if (l_ok) b_mut += (-b_cl * coef_l / mh_l) * g.c1h[f - 1];
    // wn = w_work + dts rw
    bar[A_w_work][Q(f)] += bw;
    const double b_rw = bw * dts;
    bar[A_rw_tend][Q(f)] += b_rw;
    if (nl.w_damping == 1) {
      const double mass_f = MF(f);
      const bool mf_ok = fabs(mass_f) > 1.0e-12;
      const double safe_mf = mf_ok ? mass_f : 1.0e-12;
      const double qv = WW(f) / safe_mf;
      const double z = (qv * g.rdnw[f]) * dts;
      const double vert_cfl = fabs(z);
      if (vert_cfl > kWBeta) {
        const double sA = (-sgn0(WQ(f))) * kWAlpha;
        const double b_vc = b_rw * mass_f * sA;
        double b_mass_f = b_rw * (sA * (vert_cfl - kWBeta));
        const double b_z = b_vc * sgn0(z);
        const double b_q = b_z * dts * g.rdnw[f];
        bar[A_ww][Q(f)] += b_q / safe_mf;
        if (mf_ok) b_mass_f += -b_q * qv / safe_mf;
        b_mut += b_mass_f * g.c1f[f];
      }
    }
  }
  // rhs after the mass scaling (rhs[nz] = 0 under the lid)
  for (int f = 1; f <= nz; ++f) {
    if (f == nz && nl.top_lid) continue;
    const double b = S(8, f);
    bar[A_ph_work][Q(f)] += b;
    const double mass_f = MF(f);
    const double b_ra = b * msfty / mass_f;
    b_mut += (-b * (msfty * S(1, f)) / (mass_f * mass_f)) * g.c1f[f];
    // rhs before the scaling: dts (ph_tend + c05geps w_work) - dts (fnm wd_f1 + fnp wd_f)
    bar[A_ph_tend][Q(f)] += b_ra * dts;
    bar[A_w_work][Q(f)] += b_ra * dts * c05geps;
    if (f < nz) {
      const double b_adv = -b_ra * dts;
      const double b_wd_f1 = b_adv * g.fnm[f], b_wd_f = b_adv * g.fnp[f];
      const double s1 = 0.5 * (WW(f + 1) + WW(f)), d1 = PH1T(f + 1) - PH1T(f);
      const double s0 = 0.5 * (WW(f) + WW(f - 1)), d0 = PH1T(f) - PH1T(f - 1);
      const double bw1 = b_wd_f1 * g.rdnw[f] * d1 * 0.5, bw0 = b_wd_f * g.rdnw[f - 1] * d0 * 0.5;
      bar[A_ww][Q(f + 1)] += bw1;
      bar[A_ww][Q(f)] += bw1 + bw0;
      bar[A_ww][Q(f - 1)] += bw0;
      const double bp1 = b_wd_f1 * (s1 * g.rdnw[f]), bp0 = b_wd_f * (s0 * g.rdnw[f - 1]);
      bar[A_ph0][Q(f + 1)] += bp1;
      acc(&bar[A_phb][Q(f + 1)], bp1);
      bar[A_ph0][Q(f)] += -bp1 + bp0;
      acc(&bar[A_phb][Q(f)], -bp1 + bp0);
      bar[A_ph0][Q(f - 1)] += -bp0;
      acc(&bar[A_phb][Q(f - 1)], -bp0);
    }
  }
  // t_2ave
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k);
    const double bt = S(9, k);
    const double mass_h = g.c1h[k] * muts + g.c2h[k];
    const double ttr = kT0 + (val[A_theta][q] - kT0);
    const double D = mass_h * ttr;
    const double b_N = bt / D, b_D = -bt * S(0, k) / D;
    bar[A_th_work][q] += b_N * 0.5 * eps_p;
    bar[A_t2ave][q] += b_N * 0.5 * eps_m;
    b_muave += b_N * g.c1h[k] * kT0;
    b_muts += (b_D * ttr) * g.c1h[k];
    bar[A_theta][q] += b_D * mass_h;
  }
  bar[A_mut][c] += b_mut;
  bar[A_muts][c] += b_muts;
  bar[A_muave][c] += b_muave;
}

// ---------------------------------------------------------------------------
// advance_mu_t (periodic) with the theta update on column (j, i): the
// column divergence and mu, muts, muave, mudf, the omega recurrence ww, and
// theta's flux-form update; then ww = ww_new and mudf = mudf_new.
// ---------------------------------------------------------------------------
template <class Acc>
__host__ __device__ inline void adj_mu_theta_column(const DGrid& g, const DNamelist& nl, double dts, const View& V,
                                                    int j, int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const size_t c = i2(g, j, i);
  const double* const* val = V.val;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double msftx = g.msftx[c], msfty = g.msfty[c];
  const double rinv_n = 1.0 / g.msfvx[i2v(g, j + 1, i)], rinv_s = 1.0 / g.msfvx[i2v(g, j, i)];
  const double msfuy_e = g.msfuy[i2u(g, j, i + 1)], msfuy_w = g.msfuy[i2u(g, j, i)];
  const double muv_n = val[A_muv][i2v(g, j + 1, i)], muv_s = val[A_muv][i2v(g, j, i)];
  const double muu_e = val[A_muu][i2u(g, j, i + 1)], muu_w = val[A_muu][i2u(g, j, i)];
  const int ip = wrap_h(i + 1, nx), im1 = wrap_h(i - 1, nx), jp = wrap_h(j + 1, ny), jm = wrap_h(j - 1, ny);
  auto T1 = [=](int k, int jj, int ii) { return val[A_theta][im(g, k, jj, ii)] - kT0; };

  // ---- forward recompute: divergence, mu, ww ----
  double dmdt = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double c1 = g.c1h[k], c2 = g.c2h[k];
    const double v_n = val[A_v_work][iv(g, k, j + 1, i)] + ((c1 * muv_n + c2) * val[A_v][iv(g, k, j + 1, i)]) * rinv_n;
    const double v_s = val[A_v_work][iv(g, k, j, i)] + ((c1 * muv_s + c2) * val[A_v][iv(g, k, j, i)]) * rinv_s;
    const double u_e = val[A_u_work][iu(g, k, j, i + 1)] + ((c1 * muu_e + c2) * val[A_u][iu(g, k, j, i + 1)]) / msfuy_e;
    const double u_w = val[A_u_work][iu(g, k, j, i)] + ((c1 * muu_w + c2) * val[A_u][iu(g, k, j, i)]) / msfuy_w;
    const double d = (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
    dmdt = (k == 0) ? g.dnw[0] * d : dmdt + g.dnw[k] * d;
  }
  const double muts = val[A_muts][c], mut = val[A_mut][c];
  // ww_new, from the recurrence (recomputed level by level in the reverse pass)
  auto d_at = [=](int k) {
    const double c1 = g.c1h[k], c2 = g.c2h[k];
    const double v_n = val[A_v_work][iv(g, k, j + 1, i)] + ((c1 * muv_n + c2) * val[A_v][iv(g, k, j + 1, i)]) * rinv_n;
    const double v_s = val[A_v_work][iv(g, k, j, i)] + ((c1 * muv_s + c2) * val[A_v][iv(g, k, j, i)]) * rinv_s;
    const double u_e = val[A_u_work][iu(g, k, j, i + 1)] + ((c1 * muu_e + c2) * val[A_u][iu(g, k, j, i + 1)]) / msfuy_e;
    const double u_w = val[A_u_work][iu(g, k, j, i)] + ((c1 * muu_w + c2) * val[A_u][iu(g, k, j, i)]) / msfuy_w;
    return (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
  };
  // ww_new[f] for the theta fluxes: wk chain from ww[0]
  auto ww_new_at = [=](int f) {
    if (f == 0) return val[A_ww][Q(0)] - val[A_rom][Q(0)];
    if (f == nz) return val[A_ww][Q(nz)];
    double wk = val[A_ww][Q(0)];
    for (int kk = 1; kk <= f; ++kk) {
      const int k = kk - 1;
      const double incr = (g.dnw[kk - 1] * ((g.c1h[k] * dmdt + d_at(kk - 1)) + g.c1h[k] * 0.0)) / msfty;
      wk = wk - incr;
    }
    return wk - val[A_rom][Q(f)];
  };

  // ---- reverse ----
  // ww = ww_new, mudf = mudf_new (copies); the old mudf is not read
  double* const wwn_bar = bar[A_ww_new];
  for (int f = 0; f <= nz; ++f) {
    wwn_bar[Q(f)] += bar[A_ww][Q(f)];
    bar[A_ww][Q(f)] = 0.0;
  }
  const double b_mudf_new = bar[A_mudf_new][c] + bar[A_mudf][c];
  bar[A_mudf][c] = 0.0;
  bar[A_mudf_new][c] = 0.0;
  // theta update, k = nz-1 .. 0 (th_work is read and overwritten in place)
  double carry = 0.0;  // -bar(wd_lo) of level k+1 = bar contribution to wd_hi(k)
  for (int k = nz - 1; k >= 0; --k) {
    const size_t q = Q(k);
    const double b_thn = bar[A_th_work][q];
    bar[A_tht][q] += b_thn * (msfty * dts);
    const double b_tend = -b_thn * (dts * msfty);
    const double b_vf = b_tend * msftx * (0.5 * g.rdy), b_uf = b_tend * msftx * (0.5 * g.rdx);
    const double b_wdd = b_tend * g.rdnw[k];
    const double th = T1(k, j, i);
    const double vwn = val[A_v_work][iv(g, k, j + 1, i)], vws = val[A_v_work][iv(g, k, j, i)];
    const double uwe = val[A_u_work][iu(g, k, j, i + 1)], uww = val[A_u_work][iu(g, k, j, i)];
    acc(&bar[A_v_work][iv(g, k, j + 1, i)], b_vf * (T1(k, jp, i) + th));
    acc(&bar[A_v_work][iv(g, k, j, i)], -b_vf * (th + T1(k, jm, i)));
    acc(&bar[A_u_work][iu(g, k, j, i + 1)], b_uf * (T1(k, j, ip) + th));
    acc(&bar[A_u_work][iu(g, k, j, i)], -b_uf * (th + T1(k, j, im1)));
    acc(&bar[A_theta][im(g, k, jp, i)], b_vf * vwn);
    acc(&bar[A_theta][im(g, k, jm, i)], -b_vf * vws);
    acc(&bar[A_theta][im(g, k, j, ip)], b_uf * uwe);
    acc(&bar[A_theta][im(g, k, j, im1)], -b_uf * uww);
    acc(&bar[A_theta][q], (b_vf * vwn - b_vf * vws) + (b_uf * uwe - b_uf * uww));
    // wd_hi(k) = ww_new[k+1] face(k+1) below the top level
    const double b_wd_hi = b_wdd + carry;
    carry = -b_wdd;
    if (k + 1 < nz) {
      const int f = k + 1;
      const double face = g.fnm[f] * T1(f, j, i) + g.fnp[f] * T1(f - 1, j, i);
      wwn_bar[Q(f)] += b_wd_hi * face;
      const double b_face = b_wd_hi * ww_new_at(f);
      acc(&bar[A_theta][Q(f)], b_face * g.fnm[f]);
      acc(&bar[A_theta][Q(f - 1)], b_face * g.fnp[f]);
    }
    // th_work's old value enters with derivative 1; b_thn stays as its bar
  }
  // ww_new recurrence
  double b_dmdt = 0.0;
  bar[A_rom][Q(0)] -= wwn_bar[Q(0)];
  bar[A_ww][Q(0)] += wwn_bar[Q(0)];
  bar[A_ww][Q(nz)] += wwn_bar[Q(nz)];
  wwn_bar[Q(0)] = 0.0;
  wwn_bar[Q(nz)] = 0.0;
  double b_wk = 0.0;
  for (int kk = nz - 1; kk >= 1; --kk) {
    const int k = kk - 1;
    b_wk += wwn_bar[Q(kk)];
    bar[A_rom][Q(kk)] -= wwn_bar[Q(kk)];
    wwn_bar[Q(kk)] = 0.0;
    const double b_inner = (-b_wk) * g.dnw[kk - 1] / msfty;
    b_dmdt += b_inner * g.c1h[k];
    bar[A_dvdxi][Q(kk - 1)] += b_inner;  // bar of d[kk-1], collected below
  }
  bar[A_ww][Q(0)] += b_wk;
  // mu, muts, muave, mudf_new
  {
    const double b_mu_n = bar[A_mu][c], b_muts_n = bar[A_muts][c], b_muave_n = bar[A_muave][c];
    const double b_mwn = (b_mu_n + b_muts_n) + 0.5 * eps_p * b_muave_n;
    double b_mwo = 0.5 * eps_m * b_muave_n;
    double b_mut = b_muts_n;
    const double b_mu_tend = b_mudf_new + dts * b_mwn;
    b_mwo += b_mwn;
    const double b_mu_old = b_mu_n;
    b_mwo -= b_mu_n;
    bar[A_mu][c] = b_mu_old;
    bar[A_muts][c] = b_mwo;
    bar[A_muave][c] = 0.0;
    b_mut -= b_mwo;
    bar[A_mut][c] += b_mut;
    b_dmdt += b_mu_tend;
  }
  // dmdt = sum dnw[k] d[k]; d (dvdxi) is written, not read, so its bar before is 0
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k);
    const double b_d = bar[A_dvdxi][q] + b_dmdt * g.dnw[k];
    bar[A_dvdxi][q] = 0.0;
    const double b_in = b_d * (msftx * msfty);
    const double b_vn = b_in * g.rdy, b_vs = -b_in * g.rdy, b_ue = b_in * g.rdx, b_uw = -b_in * g.rdx;
    const double c1 = g.c1h[k], c2 = g.c2h[k];
    const double vn = val[A_v][iv(g, k, j + 1, i)], vs = val[A_v][iv(g, k, j, i)];
    const double ue = val[A_u][iu(g, k, j, i + 1)], uw = val[A_u][iu(g, k, j, i)];
    acc(&bar[A_v_work][iv(g, k, j + 1, i)], b_vn);
    acc(&bar[A_v][iv(g, k, j + 1, i)], b_vn * ((c1 * muv_n + c2) * rinv_n));
    acc(&bar[A_muv][i2v(g, j + 1, i)], b_vn * rinv_n * vn * c1);
    acc(&bar[A_v_work][iv(g, k, j, i)], b_vs);
    acc(&bar[A_v][iv(g, k, j, i)], b_vs * ((c1 * muv_s + c2) * rinv_s));
    acc(&bar[A_muv][i2v(g, j, i)], b_vs * rinv_s * vs * c1);
    acc(&bar[A_u_work][iu(g, k, j, i + 1)], b_ue);
    acc(&bar[A_u][iu(g, k, j, i + 1)], b_ue * ((c1 * muu_e + c2) / msfuy_e));
    acc(&bar[A_muu][i2u(g, j, i + 1)], b_ue * ue * c1 / msfuy_e);
    acc(&bar[A_u_work][iu(g, k, j, i)], b_uw);
    acc(&bar[A_u][iu(g, k, j, i)], b_uw * ((c1 * muu_w + c2) / msfuy_w));
    acc(&bar[A_muu][i2u(g, j, i)], b_uw * uw * c1 / msfuy_w);
  }
}

// ---------------------------------------------------------------------------
// advance_uv on one face column: face (j, i) of u (ydir = false, i in
// 0..nx) or of v (ydir = true, j in 0..ny).  The pressure-gradient pair is
// edge-clamped (cells il/ir or js/jn).
// ---------------------------------------------------------------------------
template <class Acc>
__host__ __device__ inline void adj_uv_face(const DGrid& g, double dts, const View& V, bool ydir, int j, int i,
                                            Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  int jl, il, jr, ir;
  size_t f2;
  double msf, rd, mudf_scale;
  if (!ydir) {
    jl = jr = j;
    il = clamp_h(i - 1, 0, nx - 1);
    ir = clamp_h(i, 0, nx - 1);
    f2 = i2u(g, j, i);
    msf = g.msfux[f2] / g.msfuy[f2];
    rd = g.rdx;
    mudf_scale = (-kEmdiv * g.dx) / g.msfuy[f2];  // d mudf_xy / d (mudf_r - mudf_l)
  } else {
    il = ir = i;
    jl = clamp_h(j - 1, 0, ny - 1);
    jr = clamp_h(j, 0, ny - 1);
    f2 = i2v(g, j, i);
    msf = g.msfvy[f2] / g.msfvx[f2];
    rd = g.rdy;
    mudf_scale = (-kEmdiv * g.dy) * (1.0 / g.msfvx[f2]);
  }
  const size_t cl = i2(g, jl, il), cr = i2(g, jr, ir);
  const int w_arr = ydir ? A_v_work : A_u_work, t_arr = ydir ? A_vt : A_ut, m_arr = ydir ? A_muv : A_muu;
  auto F = [=](int k) { return ydir ? iv(g, k, j, i) : iu(g, k, j, i); };
  auto L = [=](int k) { return im(g, k, jl, il); };
  auto R = [=](int k) { return im(g, k, jr, ir); };
  const double mface = val[m_arr][f2];
  double b_mface = 0.0, b_mudf_xy = 0.0, b_muw = 0.0;
  double carry = 0.0;  // bar of dpn[k] from bracket k-1
  // the dpn bar of level k goes to the pair sums of p
  auto dpn_bar = [&](int k, double b) {
    const double h = 0.5 * b;
    auto add_ps = [&](int kk, double v) {
      acc(&bar[A_p][L(kk)], v);
      acc(&bar[A_p][R(kk)], v);
    };
    if (k == 0) {
      add_ps(0, h * g.cf1);
      add_ps(1, h * g.cf2);
      add_ps(2, h * g.cf3);
    } else if (k == nz) {
      add_ps(nz - 1, h * g.cf1);
      add_ps(nz - 2, h * g.cf2);
      add_ps(nz - 3, h * g.cf3);
    } else {
      add_ps(k, h * g.fnm[k]);
      add_ps(k - 1, h * g.fnp[k]);
    }
  };
  for (int k = 0; k < nz; ++k) {
    const size_t q = F(k);
    const double b_un = bar[w_arr][q];  // work array: read and overwritten in place
    bar[t_arr][q] += b_un * dts;
    b_mudf_xy += b_un * g.c1h[k];
    const double b_dpx = -b_un * (dts * 1.0);
    // dpx = (C mass) T + G bracket
    const double mass = g.c1h[k] * mface + g.c2h[k];
    const double C = (msf * 0.5) * rd;
    const double php_d = val[A_php][R(k)] - val[A_php][L(k)];
    const double Gk = (msf * rd) * php_d;
    const double ph_term = (val[A_ph_work][R(k + 1)] - val[A_ph_work][L(k + 1)]) +
                           (val[A_ph_work][R(k)] - val[A_ph_work][L(k)]);
    const double alt_s = val[A_alt][L(k)] + val[A_alt][R(k)];
    const double p_d = val[A_p][R(k)] - val[A_p][L(k)];
    const double al_s = val[A_al][L(k)] + val[A_al][R(k)];
    const double pb_d = val[A_pb][R(k)] - val[A_pb][L(k)];
    const double T = (ph_term + alt_s * p_d) + al_s * pb_d;
    // bracket needs dpn, recomputed from p
    auto ps = [=](int kk) { return val[A_p][L(kk)] + val[A_p][R(kk)]; };
    auto dpn = [=](int kk) {
      if (kk == 0) return 0.5 * ((g.cf1 * ps(0) + g.cf2 * ps(1)) + g.cf3 * ps(2));
      if (kk == nz) return 0.5 * ((g.cf1 * ps(nz - 1) + g.cf2 * ps(nz - 2)) + g.cf3 * ps(nz - 3));
      return 0.5 * (g.fnm[kk] * ps(kk) + g.fnp[kk] * ps(kk - 1));
    };
    const double muw_s = (val[A_muts][cl] - val[A_mut][cl]) + (val[A_muts][cr] - val[A_mut][cr]);
    const double bracket = g.rdnw[k] * (dpn(k + 1) - dpn(k)) - 0.5 * (g.c1h[k] * muw_s);
    const double b_G = b_dpx * bracket, b_br = b_dpx * Gk;
    acc(&bar[A_php][R(k)], b_G * (msf * rd));
    acc(&bar[A_php][L(k)], -b_G * (msf * rd));
    // dpn[k] is complete once bracket k is in: it gets -rdnw[k] b_br here and
    // +rdnw[k-1] b_br(k-1) from the level below
    dpn_bar(k, carry - b_br * g.rdnw[k]);
    carry = b_br * g.rdnw[k];
    b_muw += -b_br * 0.5 * g.c1h[k];
    b_mface += (b_dpx * C * T) * g.c1h[k];
    const double b_T = b_dpx * (C * mass);
    acc(&bar[A_ph_work][R(k + 1)], b_T);
    acc(&bar[A_ph_work][L(k + 1)], -b_T);
    acc(&bar[A_ph_work][R(k)], b_T);
    acc(&bar[A_ph_work][L(k)], -b_T);
    acc(&bar[A_alt][L(k)], b_T * p_d);
    acc(&bar[A_alt][R(k)], b_T * p_d);
    acc(&bar[A_p][R(k)], b_T * alt_s);
    acc(&bar[A_p][L(k)], -b_T * alt_s);
    acc(&bar[A_al][L(k)], b_T * pb_d);
    acc(&bar[A_al][R(k)], b_T * pb_d);
    acc(&bar[A_pb][R(k)], b_T * al_s);
    acc(&bar[A_pb][L(k)], -b_T * al_s);
  }
  dpn_bar(nz, carry);
  bar[m_arr][f2] += b_mface;
  acc(&bar[A_muts][cl], b_muw);
  acc(&bar[A_mut][cl], -b_muw);
  acc(&bar[A_muts][cr], b_muw);
  acc(&bar[A_mut][cr], -b_muw);
  acc(&bar[A_mudf][cr], b_mudf_xy * mudf_scale);
  acc(&bar[A_mudf][cl], -b_mudf_xy * mudf_scale);
}

}  // namespace adj
}  // namespace wd
