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

  // ---- forward recompute: the column divergence d (scratch plane 0), its
  // sum dmdt, and ww_new (scratch plane 1) ----
  const size_t P = (size_t)(nz + 1) * ny * nx;
  double* const d_s = V.scr;
  double* const wwn_s = V.scr + P;
  double dmdt = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double c1 = g.c1h[k], c2 = g.c2h[k];
    const double v_n = val[A_v_work][iv(g, k, j + 1, i)] + ((c1 * muv_n + c2) * val[A_v][iv(g, k, j + 1, i)]) * rinv_n;
    const double v_s = val[A_v_work][iv(g, k, j, i)] + ((c1 * muv_s + c2) * val[A_v][iv(g, k, j, i)]) * rinv_s;
    const double u_e = val[A_u_work][iu(g, k, j, i + 1)] + ((c1 * muu_e + c2) * val[A_u][iu(g, k, j, i + 1)]) / msfuy_e;
    const double u_w = val[A_u_work][iu(g, k, j, i)] + ((c1 * muu_w + c2) * val[A_u][iu(g, k, j, i)]) / msfuy_w;
    const double d = (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
    d_s[Q(k)] = d;
    dmdt = (k == 0) ? g.dnw[0] * d : dmdt + g.dnw[k] * d;
  }
  const double muts = val[A_muts][c], mut = val[A_mut][c];
  (void)muts;
  (void)mut;
  {
    double wk = val[A_ww][Q(0)];
    wwn_s[Q(0)] = wk - val[A_rom][Q(0)];
    for (int kk = 1; kk < nz; ++kk) {
      const int k = kk - 1;
      const double incr = (g.dnw[kk - 1] * ((g.c1h[k] * dmdt + d_s[Q(kk - 1)]) + g.c1h[k] * 0.0)) / msfty;
      wk = wk - incr;
      wwn_s[Q(kk)] = wk - val[A_rom][Q(kk)];
    }
    wwn_s[Q(nz)] = val[A_ww][Q(nz)];
  }
  auto ww_new_at = [=](int f) { return wwn_s[Q(f)]; };

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

// ---------------------------------------------------------------------------
// Step origin (operational_mode.py:3328-3329): u0, v0, w0, th0, ph0, mu0 are
// copies of the state the large step starts from.  Element q of one pair.
// ---------------------------------------------------------------------------
__host__ __device__ inline void adj_origin_elem(const View& V, int src, int dst, size_t q) {
  V.bar[src][q] += V.bar[dst][q];
  V.bar[dst][q] = 0.0;
}

// The equation-of-state pressure of the refresh and its partial derivatives
// (dynamics/acoustic_wrf.py:198-211).
struct EosP {
  double p, dtheta, dalt;
};
__host__ __device__ inline EosP eos_p_from_theta_alt(double theta, double alt) {
  const bool alt_ok = alt > kMinAlt;
  const double as = alt_ok ? alt : kMinAlt;
  const double den = kP0 * as;
  const double arg = ((kRd * theta) * 1.0) / den;
  const bool arg_ok = arg > 1.0e-12;
  const double a2 = arg_ok ? arg : 1.0e-12;
  const double pw = wd_pow(a2, kCpovcv);
  const double dp_da2 = kP0 * ((kCpovcv * pw) / a2);
  EosP r;
  r.p = kP0 * pw;
  r.dtheta = arg_ok ? dp_da2 * (kRd * 1.0 / den) : 0.0;
  r.dalt = (arg_ok && alt_ok) ? dp_da2 * (-arg / as) : 0.0;
  return r;
}

__host__ __device__ inline double safe_den(double ms) {
  return fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12);
}

// ---------------------------------------------------------------------------
// small_step_finish (core/small_step_finish.py:22-101) and the p refresh
// (operational_mode.py:2169-2231, acoustic_wrf.py:360-438) on column (j, i):
// w, ph, theta, p and mu.  u and v are adj_finish_face.
// ---------------------------------------------------------------------------
__host__ __device__ inline void adj_finish_column(const DGrid& g, const View& V, int j, int i) {
  const int nz = g.nz;
  const size_t c = i2(g, j, i);
  const double* const* val = V.val;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double mut = val[A_mut][c], muts = val[A_muts][c], mub = val[A_mub][c];
  const double msfty = g.msfty[c];
  // values after finish that the refresh reads
  const double mu_pert_n = val[A_mu][c];
  auto ph_pert_n = [=](int k) { return val[A_ph_work][Q(k)] + val[A_ph_pert][Q(k)]; };
  auto ph_total_n = [=](int k) {
    return (val[A_ph_total][Q(k)] - val[A_ph_pert][Q(k)]) + ph_pert_n(k);
  };
  auto theta_n = [=](int k) {
    const double ms = g.c1h[k] * muts + g.c2h[k], mc = g.c1h[k] * mut + g.c2h[k];
    return (val[A_th_work][Q(k)] + (val[A_theta][Q(k)] - kT0) * mc) / safe_den(ms) + kT0;
  };
  double b_mub = 0.0, b_mut = 0.0, b_muts = 0.0, b_mupn = 0.0;
  // ---- refresh reverse (k independent but for the ph pair) ----
  // bars of the finished ph_pert / ph_total / theta, collected per level
  const double muts_r = mub + mu_pert_n;
  double b_php_up = 0.0, b_pht_up = 0.0;  // from level k to the finished ph at k+1
  for (int k = nz - 1; k >= 0; --k) {
    const size_t q = Q(k), q1 = Q(k + 1);
    const double phb_k = ph_total_n(k) - ph_pert_n(k), phb_k1 = ph_total_n(k + 1) - ph_pert_n(k + 1);
    const double mhb = g.c1h[k] * mub + g.c2h[k];
    const double dalb = g.dnw[k] * mhb;
    const bool da_ok = fabs(dalb) > 1.0e-12;
    const double safe_da = da_ok ? dalb : 1.0e-12;
    const double alb = -(phb_k1 - phb_k) / safe_da;
    const double mass_w = g.c1h[k] * muts_r + g.c2h[k];
    const bool mw_ok = fabs(mass_w) > 1.0e-12;
    const double safe_mass = mw_ok ? mass_w : 1.0e-12;
    const double mu_term = (alb * g.c1h[k]) * mu_pert_n;
    const double geo = g.rdnw[k] * (ph_pert_n(k + 1) - ph_pert_n(k));
    const double al = -(mu_term + geo) / safe_mass;
    const EosP e = eos_p_from_theta_alt(theta_n(k), al + alb);
    // p_total = (p_total_f - p_pert_f) + p_pert; p_pert = ptot - pb
    const double b_pt = bar[A_p_total][q], b_pp = bar[A_p_pert][q];
    const double b_pr = b_pp + b_pt;
    // bars of the finished p_total / p_pert, consumed by the finish below
    bar[A_p_total][q] = b_pt;
    bar[A_p_pert][q] = -b_pt;
    bar[A_pb][q] += -b_pr;
    const double b_theta_n = b_pr * e.dtheta;
    const double b_alt = b_pr * e.dalt;
    const double b_num = -b_alt / safe_mass;
    const double b_safe_mass = b_alt * (mu_term + geo) / (safe_mass * safe_mass);
    double b_alb = b_alt + b_num * g.c1h[k] * mu_pert_n;
    b_mupn += b_num * (alb * g.c1h[k]);
    const double b_geo = b_num * g.rdnw[k];
    if (mw_ok) {
      b_mub += b_safe_mass * g.c1h[k];
      b_mupn += b_safe_mass * g.c1h[k];
    }
    const double b_dphb = -b_alb / safe_da;
    if (da_ok) b_mub += (b_alb * (phb_k1 - phb_k) / (safe_da * safe_da)) * g.dnw[k] * g.c1h[k];
    // phb = ph_total_n - ph_pert_n at k+1 (+b_dphb) and k (-b_dphb); geo on ph_pert_n
    b_pht_up += b_dphb;
    b_php_up += -b_dphb + b_geo;
    V.scr[q1] = b_pht_up;  // finished ph_total bar at k+1 (scratch plane 0)
    V.scr[(size_t)(nz + 1) * g.ny * g.nx + q1] = b_php_up;  // finished ph_pert bar at k+1 (plane 1)
    b_pht_up = -b_dphb;
    b_php_up = b_dphb - b_geo;
    V.scr[(size_t)2 * (nz + 1) * g.ny * g.nx + q] = b_theta_n;  // plane 2
  }
  V.scr[Q(0)] = b_pht_up;
  V.scr[(size_t)(nz + 1) * g.ny * g.nx + Q(0)] = b_php_up;
  // ---- finish reverse ----
  // mu: mu_pert = W.mu, mu_total = (mu_total - mu_pert) + W.mu
  {
    const double b_mutot = bar[A_mu_total][c];
    b_mupn += bar[A_mu_pert][c];
    bar[A_mu][c] += b_mupn + b_mutot;
    bar[A_mu_total][c] = b_mutot;
    bar[A_mu_pert][c] = -b_mutot;
    bar[A_mub][c] += b_mub;
  }
  const size_t P = (size_t)(nz + 1) * g.ny * g.nx;
  for (int k = 0; k <= nz; ++k) {
    const size_t q = Q(k);
    // ph: ph_pert_n = ph_work + ph_pert; ph_total_n = (ph_total - ph_pert) + ph_pert_n
    const double b_pht_n = bar[A_ph_total][q] + V.scr[q];
    const double b_php_n = bar[A_ph_pert][q] + V.scr[P + q];
    bar[A_ph_work][q] += b_php_n + b_pht_n;
    bar[A_ph_pert][q] = b_php_n;
    bar[A_ph_total][q] = b_pht_n;
    // w = (msfty w_work + w mc) / den
    const double ms = g.c1f[k] * muts + g.c2f[k], mc = g.c1f[k] * mut + g.c2f[k];
    const double den = safe_den(ms);
    const double w_old = val[A_w][q];
    const double w_n = (msfty * val[A_w_work][q] + w_old * mc) / den;
    const double b_wn = bar[A_w][q];
    bar[A_w_work][q] += b_wn * msfty / den;
    bar[A_w][q] = b_wn * mc / den;
    b_mut += (b_wn * w_old / den) * g.c1f[k];
    if (fabs(ms) > 1.0e-12) b_muts += (-b_wn * w_n / den) * g.c1f[k];
  }
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k);
    // theta_n = (th_work + (theta - T0) mc) / den + T0
    const double b_thn = bar[A_theta][q] + V.scr[2 * P + q];
    const double ms = g.c1h[k] * muts + g.c2h[k], mc = g.c1h[k] * mut + g.c2h[k];
    const double den = safe_den(ms);
    const double t_save = val[A_theta][q] - kT0;
    const double tp = (val[A_th_work][q] + t_save * mc) / den;
    bar[A_th_work][q] += b_thn / den;
    bar[A_theta][q] = b_thn * mc / den;
    b_mut += (b_thn * t_save / den) * g.c1h[k];
    if (fabs(ms) > 1.0e-12) b_muts += (-b_thn * tp / den) * g.c1h[k];
    // p_total_f = (p_total - p_pert) + W.p; p_pert_f = W.p
    const double b_ptf = bar[A_p_total][q], b_ppf = bar[A_p_pert][q];
    bar[A_p][q] += b_ptf + b_ppf;
    bar[A_p_total][q] = b_ptf;
    bar[A_p_pert][q] = -b_ptf;
  }
  bar[A_mut][c] += b_mut;
  bar[A_muts][c] += b_muts;
}

// u face (j, i) (ydir false) or v face (ydir true) of small_step_finish.
__host__ __device__ inline void adj_finish_face(const DGrid& g, const View& V, bool ydir, int j, int i) {
  const int nz = g.nz;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const size_t f2 = ydir ? i2v(g, j, i) : i2u(g, j, i);
  const int st = ydir ? A_v : A_u, wk = ydir ? A_v_work : A_u_work;
  const int ms_arr = ydir ? A_muvs : A_muus, mc_arr = ydir ? A_muv : A_muu;
  const double msf = ydir ? g.msfvx[f2] : g.msfuy[f2];
  const double mfs = val[ms_arr][f2], mfc = val[mc_arr][f2];
  double b_ms = 0.0, b_mc = 0.0;
  for (int k = 0; k < nz; ++k) {
    const size_t q = ydir ? iv(g, k, j, i) : iu(g, k, j, i);
    const double ms = g.c1h[k] * mfs + g.c2h[k], mc = g.c1h[k] * mfc + g.c2h[k];
    const double den = safe_den(ms);
    const double x_old = val[st][q];
    const double x_n = (msf * val[wk][q] + x_old * mc) / den;
    const double b = bar[st][q];
    bar[wk][q] += b * msf / den;
    bar[st][q] = b * mc / den;
    b_mc += (b * x_old / den) * g.c1h[k];
    if (fabs(ms) > 1.0e-12) b_ms += (-b * x_n / den) * g.c1h[k];
  }
  bar[mc_arr][f2] += b_mc;
  bar[ms_arr][f2] += b_ms;
}

// ===========================================================================
// Stage setup (operational_mode.py:2429-3093, small_step_prep.py:216-394,
// calc_p_rho.py:48-139, advance_w.py:209-290, rhs_ph.py:175-398,
// acoustic_wrf.py:783-852), reversed in phases.  Each phase finishes on the
// whole grid before the next starts, because phases scatter into bars that
// later phases consume:
//   S1 column  carried-state init, calc_coef_w, rhs_ph, rw_tend (+ wt and
//              the curvature term), calc_p_rho step 0
//   S2 column  php, then w_work / ph_work / phb, th_work / pb / alt / c2a
//   S3 face    u_work / v_work
//   S4 column  the face averages of mut and muts0, then the 2-D prep
//   S5 face    large-step PGF on u and v
//   S6 column  al_ls / php_ls (_absolute_diagnostics) and the EOS alt
//   S7         flux advection of theta - T0, u, v, w (adj_advect_*)
//   S8 column  rom and the column divergence (couple_velocities)
//   S9 point   ru, rv (couple_velocities)
// Every array the setup writes has a zero bar before it (none is read
// before it is written).
// ===========================================================================

__host__ __device__ inline double mh_at(const DGrid& g, int k, double mu) { return g.c1h[k] * mu + g.c2h[k]; }
__host__ __device__ inline double mf_at(const DGrid& g, int k, double mu) { return g.c1f[k] * mu + g.c2f[k]; }

template <class Acc>
__host__ __device__ inline void adj_setup_s1_column(const DGrid& g, const DNamelist& nl, double dts, const View& V,
                                                    int j, int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const size_t c = i2(g, j, i);
  const double* const* val = V.val;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double mut = val[A_mut][c], mub = val[A_mub][c];
  const double msfty = g.msfty[c], msftx = g.msftx[c];
  const double msft_inv = 1.0 / msfty;
  double b_mut = 0.0, b_mub = 0.0;
  // --- carried init: muts = muts0, muave = 0, ww = rom, t2ave = 0 ---
  bar[A_muts0][c] += bar[A_muts][c];
  bar[A_muts][c] = 0.0;
  bar[A_muave][c] = 0.0;
  for (int f = 0; f <= nz; ++f) {
    acc(&bar[A_rom][Q(f)], bar[A_ww][Q(f)]);
    bar[A_ww][Q(f)] = 0.0;
  }
  for (int k = 0; k < nz; ++k) bar[A_t2ave][Q(k)] = 0.0;
  // --- calc_coef_w (top_lid: a[nz] = 0) ---
  {
    double cof = (0.5 * dts * kG * (1.0 + nl.epssm));
    cof = cof * cof;
    auto MH = [=](int k) { return mh_at(g, k, mut); };
    auto MF = [=](int k) { return mf_at(g, k, mut); };
    auto C2A = [=](int k) { return val[A_c2a][Q(k)]; };
    auto cqw = [=](int k) { return (k >= 1 && k < nz) ? 1.0 : 0.0; };
    // forward recompute of a, alpha, gamma into scratch planes 0, 1, 2
    const size_t P = (size_t)(nz + 1) * ny * nx;
    double* sa = V.scr;
    double* sal = V.scr + P;
    double* sg = V.scr + 2 * P;
    for (int k = 0; k <= nz; ++k) {
      sa[Q(k)] = 0.0;
      sal[Q(k)] = 1.0;
      sg[Q(k)] = 0.0;
    }
    const double r2 = g.rdnw[nz - 1] * g.rdnw[nz - 1];
    for (int kk = 2; kk < nz; ++kk) {
      const int k = kk - 1;
      sa[Q(kk)] = (((((-cqw(kk)) * cof) * g.rdn[kk]) * g.rdnw[kk - 1]) * C2A(kk - 1)) / (MH(k) * MF(k));
    }
    for (int k = 1; k < nz; ++k) {
      const double du = MH(k) * MF(k), dl = MH(k - 1) * MF(k), dc = MH(k) * MF(k + 1);
      const double b = 1.0 + ((cqw(k) * cof) * g.rdn[k]) * ((g.rdnw[k] * C2A(k)) / du + (g.rdnw[k - 1] * C2A(k - 1)) / dl);
      const double cc = (((((-cqw(k)) * cof) * g.rdn[k]) * g.rdnw[k]) * C2A(k)) / dc;
      const double al = 1.0 / (b - sa[Q(k)] * sg[Q(k - 1)]);
      sal[Q(k)] = al;
      sg[Q(k)] = cc * al;
    }
    const double tdb = MH(nz - 1) * MF(nz);
    const double K2 = (2.0 * cof) * r2;
    const double b_top = 1.0 + (K2 * C2A(nz - 1)) / tdb;
    sal[Q(nz)] = 1.0 / (b_top - sa[Q(nz)] * sg[Q(nz - 1)]);
    // reverse
    auto dmut_prod = [&](double b, int kh, int kf) {  // d(MH(kh) MF(kf)) / d mut
      b_mut += b * (MF(kf) * g.c1h[kh] + MH(kh) * g.c1f[kf]);
    };
    double* ba = bar[A_ca];
    double* bal = bar[A_calpha];
    double* bg = bar[A_cgamma];
    {
      const double an = sal[Q(nz)];
      const double bD = -bal[Q(nz)] * an * an;
      ba[Q(nz)] += -bD * sg[Q(nz - 1)];
      bg[Q(nz - 1)] += -bD * sa[Q(nz)];
      bar[A_c2a][Q(nz - 1)] += bD * K2 / tdb;
      dmut_prod(-bD * (K2 * C2A(nz - 1)) / (tdb * tdb), nz - 1, nz);
    }
    for (int k = nz - 1; k >= 1; --k) {
      const double du = MH(k) * MF(k), dl = MH(k - 1) * MF(k), dc = MH(k) * MF(k + 1);
      const double al = sal[Q(k)];
      const double Kc = ((((-cqw(k)) * cof) * g.rdn[k]) * g.rdnw[k]);
      const double cc = (Kc * C2A(k)) / dc;
      const double b_c = bg[Q(k)] * al;
      const double b_al = bg[Q(k)] * cc + bal[Q(k)];
      const double bD = -b_al * al * al;
      ba[Q(k)] += -bD * sg[Q(k - 1)];
      bg[Q(k - 1)] += -bD * sa[Q(k)];
      bar[A_c2a][Q(k)] += b_c * Kc / dc;
      dmut_prod(-b_c * cc / dc, k, k + 1);
      const double Kb = (cqw(k) * cof) * g.rdn[k];
      const double X1 = (g.rdnw[k] * C2A(k)) / du, X2 = (g.rdnw[k - 1] * C2A(k - 1)) / dl;
      const double bX = bD * Kb;
      bar[A_c2a][Q(k)] += bX * g.rdnw[k] / du;
      bar[A_c2a][Q(k - 1)] += bX * g.rdnw[k - 1] / dl;
      dmut_prod(-bX * X1 / du, k, k);
      dmut_prod(-bX * X2 / dl, k - 1, k);
    }
    for (int kk = 2; kk < nz; ++kk) {
      const int k = kk - 1;
      const double den = MH(k) * MF(k);
      const double Ka = ((((-cqw(kk)) * cof) * g.rdn[kk]) * g.rdnw[kk - 1]);
      bar[A_c2a][Q(kk - 1)] += ba[Q(kk)] * Ka / den;
      dmut_prod(-ba[Q(kk)] * sa[Q(kk)] / den, k, k);
    }
    for (int k = 0; k <= nz; ++k) {
      ba[Q(k)] = 0.0;
      bal[Q(k)] = 0.0;
      bg[Q(k)] = 0.0;
    }
  }
  // --- rhs_ph (order 2, periodic), ph_tend faces 1..nz-1 ---
  {
    const int jp = wrap_h(j + 1, ny), jm = wrap_h(j - 1, ny), ip = wrap_h(i + 1, nx), imm = wrap_h(i - 1, nx);
    auto PHT = [=](int k, int jj, int ii) { return val[A_ph_pert][im(g, k, jj, ii)] + val[A_phb][im(g, k, jj, ii)]; };
    auto bPHT = [&](int k, int jj, int ii, double b) {
      acc(&bar[A_ph_pert][im(g, k, jj, ii)], b);
      acc(&bar[A_phb][im(g, k, jj, ii)], b);
    };
    const double* rom = val[A_rom];
    const double* u = val[A_u];
    const double* v = val[A_v];
    for (int f = 1; f < nz; ++f) {
      const double bt = bar[A_ph_tend][Q(f)];
      // t = -(fnm wd_f + fnp wd_fm) + gw + Ky fy + Kx fx
      const double b_wdf = -bt * g.fnm[f], b_wdfm = -bt * g.fnp[f];
      const double s1 = 0.5 * (rom[Q(f + 1)] + rom[Q(f)]), d1 = PHT(f + 1, j, i) - PHT(f, j, i);
      const double s0 = 0.5 * (rom[Q(f)] + rom[Q(f - 1)]), d0 = PHT(f, j, i) - PHT(f - 1, j, i);
      const double br1 = b_wdf * g.rdnw[f] * d1 * 0.5, br0 = b_wdfm * g.rdnw[f - 1] * d0 * 0.5;
      acc(&bar[A_rom][Q(f + 1)], br1);
      acc(&bar[A_rom][Q(f)], br1 + br0);
      acc(&bar[A_rom][Q(f - 1)], br0);
      const double bd1 = b_wdf * (s1 * g.rdnw[f]), bd0 = b_wdfm * (s0 * g.rdnw[f - 1]);
      bPHT(f + 1, j, i, bd1);
      bPHT(f, j, i, -bd1 + bd0);
      bPHT(f - 1, j, i, -bd0);
      // gw = ((mass_f G) w) msft_inv
      const double mass_f = mf_at(g, f, mut);
      acc(&bar[A_w][Q(f)], bt * (mass_f * kG) * msft_inv);
      b_mut += (bt * val[A_w][Q(f)] * msft_inv) * kG * g.c1f[f];
      // y advection
      const double Ky = -((0.25 * g.rdy) * msft_inv), Kx = -((0.25 * g.rdx) * msft_inv);
      const double b_fy = bt * Ky, b_fx = bt * Kx;
      const double pti = PHT(f, j, i);
      const double dn_ = PHT(f, jp, i) - pti, ds_ = pti - PHT(f, jm, i);
      const double vn = v[iv(g, f, j + 1, i)] + v[iv(g, f - 1, j + 1, i)];
      const double vs = v[iv(g, f, j, i)] + v[iv(g, f - 1, j, i)];
      const double mvn = mf_at(g, f, val[A_muv][i2v(g, j + 1, i)]), mvs = mf_at(g, f, val[A_muv][i2v(g, j, i)]);
      acc(&bar[A_muv][i2v(g, j + 1, i)], b_fy * vn * dn_ * g.c1f[f]);
      acc(&bar[A_muv][i2v(g, j, i)], b_fy * vs * ds_ * g.c1f[f]);
      const double b_vn = b_fy * mvn * dn_, b_vs = b_fy * mvs * ds_;
      acc(&bar[A_v][iv(g, f, j + 1, i)], b_vn);
      acc(&bar[A_v][iv(g, f - 1, j + 1, i)], b_vn);
      acc(&bar[A_v][iv(g, f, j, i)], b_vs);
      acc(&bar[A_v][iv(g, f - 1, j, i)], b_vs);
      const double b_dn = b_fy * (mvn * vn), b_ds = b_fy * (mvs * vs);
      bPHT(f, jp, i, b_dn);
      bPHT(f, jm, i, -b_ds);
      // x advection
      const double de_ = PHT(f, j, ip) - pti, dw_ = pti - PHT(f, j, imm);
      const double ue = u[iu(g, f, j, i + 1)] + u[iu(g, f - 1, j, i + 1)];
      const double uw = u[iu(g, f, j, i)] + u[iu(g, f - 1, j, i)];
      const double mue = mf_at(g, f, val[A_muu][i2u(g, j, i + 1)]), muw = mf_at(g, f, val[A_muu][i2u(g, j, i)]);
      acc(&bar[A_muu][i2u(g, j, i + 1)], b_fx * ue * de_ * g.c1f[f]);
      acc(&bar[A_muu][i2u(g, j, i)], b_fx * uw * dw_ * g.c1f[f]);
      const double b_ue = b_fx * mue * de_, b_uw = b_fx * muw * dw_;
      acc(&bar[A_u][iu(g, f, j, i + 1)], b_ue);
      acc(&bar[A_u][iu(g, f - 1, j, i + 1)], b_ue);
      acc(&bar[A_u][iu(g, f, j, i)], b_uw);
      acc(&bar[A_u][iu(g, f - 1, j, i)], b_uw);
      const double b_de = b_fx * (mue * ue), b_dw = b_fx * (muw * uw);
      bPHT(f, j, ip, b_de);
      bPHT(f, j, imm, -b_dw);
      bPHT(f, j, i, (-b_dn + b_ds) + (-b_de + b_dw));
    }
    for (int f = 0; f <= nz; ++f) bar[A_ph_tend][Q(f)] = 0.0;
  }
  // --- rw_tend: pg_buoy_w (q = 0), + wt, + curvature ---
  {
    const double mu_p = mut - mub;
    const double* u = val[A_u];
    const double* v = val[A_v];
    const double mxy = msftx / msfty;
    double b_mu_p = 0.0;
    for (int k = 0; k <= nz; ++k) {
      const double b = bar[A_rw_tend][Q(k)];
      bar[A_wt][Q(k)] += b;
      if (k >= 1) {
        const double K = msft_inv * kG;
        if (k < nz) {
          acc(&bar[A_p_pert][Q(k)], b * K * g.rdn[k]);
          acc(&bar[A_p_pert][Q(k - 1)], -b * K * g.rdn[k]);
        } else {
          acc(&bar[A_p_pert][Q(nz - 1)], -b * K * (2.0 * g.rdnw[nz - 1]));
        }
        b_mu_p += -b * K * g.c1f[k];
      }
      if (k >= 1 && k < nz) {
        // curv = re (ru_f u_f + (mxy rv_f) v_f), level pair k (s=0), k-1 (s=1)
        double ru_m[2], u_m[2], rv_m[2], v_m[2];
        for (int s2 = 0; s2 < 2; ++s2) {
          const int kk = k - s2;
          const double mu_l = mh_at(g, kk, val[A_muu][i2u(g, j, i)]), mu_r = mh_at(g, kk, val[A_muu][i2u(g, j, i + 1)]);
          ru_m[s2] = 0.5 * ((mu_l * u[iu(g, kk, j, i)]) / g.msfuy[i2u(g, j, i)] +
                            (mu_r * u[iu(g, kk, j, i + 1)]) / g.msfuy[i2u(g, j, i + 1)]);
          u_m[s2] = 0.5 * (u[iu(g, kk, j, i)] + u[iu(g, kk, j, i + 1)]);
          const double mv_s = mh_at(g, kk, val[A_muv][i2v(g, j, i)]), mv_n = mh_at(g, kk, val[A_muv][i2v(g, j + 1, i)]);
          rv_m[s2] = 0.5 * ((mv_s * v[iv(g, kk, j, i)]) / g.msfvx[i2v(g, j, i)] +
                            (mv_n * v[iv(g, kk, j + 1, i)]) / g.msfvx[i2v(g, j + 1, i)]);
          v_m[s2] = 0.5 * (v[iv(g, kk, j, i)] + v[iv(g, kk, j + 1, i)]);
        }
        const double ru_f = g.fnm[k] * ru_m[0] + g.fnp[k] * ru_m[1];
        const double u_f = g.fnm[k] * u_m[0] + g.fnp[k] * u_m[1];
        const double rv_f = g.fnm[k] * rv_m[0] + g.fnp[k] * rv_m[1];
        const double v_f = g.fnm[k] * v_m[0] + g.fnp[k] * v_m[1];
        const double bc = b * kReradius;
        const double b_ruf = bc * u_f, b_uf = bc * ru_f, b_rvf = bc * (mxy * v_f), b_vf = bc * (mxy * rv_f);
        for (int s2 = 0; s2 < 2; ++s2) {
          const int kk = k - s2;
          const double w = s2 == 0 ? g.fnm[k] : g.fnp[k];
          const double bru = b_ruf * w * 0.5, bu = b_uf * w * 0.5, brv = b_rvf * w * 0.5, bv = b_vf * w * 0.5;
          const double ul = u[iu(g, kk, j, i)], ur = u[iu(g, kk, j, i + 1)];
          const double vs_ = v[iv(g, kk, j, i)], vn_ = v[iv(g, kk, j + 1, i)];
          const double msl = g.msfuy[i2u(g, j, i)], msr = g.msfuy[i2u(g, j, i + 1)];
          const double mss = g.msfvx[i2v(g, j, i)], msn = g.msfvx[i2v(g, j + 1, i)];
          const double mu_l = mh_at(g, kk, val[A_muu][i2u(g, j, i)]), mu_r = mh_at(g, kk, val[A_muu][i2u(g, j, i + 1)]);
          const double mv_s = mh_at(g, kk, val[A_muv][i2v(g, j, i)]), mv_n = mh_at(g, kk, val[A_muv][i2v(g, j + 1, i)]);
          acc(&bar[A_u][iu(g, kk, j, i)], bru * mu_l / msl + bu);
          acc(&bar[A_u][iu(g, kk, j, i + 1)], bru * mu_r / msr + bu);
          acc(&bar[A_muu][i2u(g, j, i)], bru * ul / msl * g.c1h[kk]);
          acc(&bar[A_muu][i2u(g, j, i + 1)], bru * ur / msr * g.c1h[kk]);
          acc(&bar[A_v][iv(g, kk, j, i)], brv * mv_s / mss + bv);
          acc(&bar[A_v][iv(g, kk, j + 1, i)], brv * mv_n / msn + bv);
          acc(&bar[A_muv][i2v(g, j, i)], brv * vs_ / mss * g.c1h[kk]);
          acc(&bar[A_muv][i2v(g, j + 1, i)], brv * vn_ / msn * g.c1h[kk]);
        }
      }
      bar[A_rw_tend][Q(k)] = 0.0;
    }
    b_mut += b_mu_p;
    b_mub -= b_mu_p;
  }
  // --- calc_p_rho step 0 (theta from the origin th0) ---
  {
    const double muts0 = val[A_muts0][c], mu_work0 = val[A_mu_work0][c];
    double b_muts0 = 0.0, b_mw0 = 0.0;
    for (int k = 0; k < nz; ++k) {
      const size_t q = Q(k), q1 = Q(k + 1);
      const double alt = val[A_alt][q], c2a = val[A_c2a][q], thw = val[A_th_work][q];
      const double mass_h = g.c1h[k] * muts0 + g.c2h[k];
      const bool mass_ok = fabs(mass_h) > 1.0e-12;
      const double safe_mass = mass_ok ? mass_h : 1.0e-12;
      const double mu_term = g.c1h[k] * mu_work0;
      const double A = alt * mu_term + g.rdnw[k] * (val[A_ph_work][q1] - val[A_ph_work][q]);
      const double al = -A / safe_mass;
      const double theta_1 = val[A_th0][q] - kT0;
      const double ttr = kT0 + theta_1;
      const bool t_ok = fabs(ttr) > 1.0e-6;
      const double safe_t = t_ok ? ttr : 1.0e-6;
      const double D = safe_mass * safe_t;
      const double X = (alt * (thw - mu_term * theta_1)) / D;
      const double b_p0 = bar[A_p][q] + bar[A_pm1][q];
      const double b_al = bar[A_al][q] - b_p0 * c2a;
      bar[A_p][q] = 0.0;
      bar[A_pm1][q] = 0.0;
      bar[A_al][q] = 0.0;
      bar[A_c2a][q] += b_p0 * (X - al);
      const double b_X = b_p0 * c2a;
      const double b_N = b_X / D, b_D = -b_X * X / D;
      double b_alt = b_N * (thw - mu_term * theta_1);
      bar[A_th_work][q] += b_N * alt;
      double b_mu_term = -b_N * alt * theta_1;
      double b_theta_1 = -b_N * alt * mu_term;
      double b_safe_mass = b_D * safe_t;
      if (t_ok) b_theta_1 += b_D * safe_mass;
      const double b_A = -b_al / safe_mass;
      b_safe_mass += b_al * A / (safe_mass * safe_mass);
      b_alt += b_A * mu_term;
      b_mu_term += b_A * alt;
      bar[A_ph_work][q1] += b_A * g.rdnw[k];
      bar[A_ph_work][q] += -b_A * g.rdnw[k];
      if (mass_ok) b_muts0 += b_safe_mass * g.c1h[k];
      b_mw0 += b_mu_term * g.c1h[k];
      bar[A_alt][q] += b_alt;
      bar[A_th0][q] += b_theta_1;
    }
    bar[A_muts0][c] += b_muts0;
    bar[A_mu_work0][c] += b_mw0;
  }
  bar[A_mut][c] += b_mut;
  bar[A_mub][c] += b_mub;
}

// S2: php, w_work / ph_work / phb (faces), th_work / pb / alt / c2a (levels).
__host__ __device__ inline void adj_setup_s2_column(const DGrid& g, const View& V, int j, int i) {
  const int nz = g.nz;
  const size_t c = i2(g, j, i);
  const double* const* val = V.val;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double mut = val[A_mut][c], muts0 = val[A_muts0][c];
  const double msfty = g.msfty[c];
  double b_mut = 0.0, b_muts0 = 0.0;
  // php = 0.5 (((phb_k + phb_k1) + ph_pert_k) + ph_pert_k1)
  for (int k = 0; k < nz; ++k) {
    const double b = 0.5 * bar[A_php][Q(k)];
    bar[A_php][Q(k)] = 0.0;
    bar[A_phb][Q(k)] += b;
    bar[A_phb][Q(k + 1)] += b;
    bar[A_ph_pert][Q(k)] += b;
    bar[A_ph_pert][Q(k + 1)] += b;
  }
  for (int k = 0; k <= nz; ++k) {
    const size_t q = Q(k);
    const double mref = mf_at(g, k, muts0), mcur = mf_at(g, k, mut);
    // w_work = (mref w0 - mcur w) / msfty
    const double bw = bar[A_w_work][q] / msfty;
    bar[A_w_work][q] = 0.0;
    bar[A_w0][q] += bw * mref;
    bar[A_w][q] += -bw * mcur;
    b_muts0 += bw * val[A_w0][q] * g.c1f[k];
    b_mut += -bw * val[A_w][q] * g.c1f[k];
    // ph_work = ph0 - ph_pert; phb = ph_total - ph_pert
    const double bph = bar[A_ph_work][q], bpb = bar[A_phb][q];
    bar[A_ph_work][q] = 0.0;
    bar[A_phb][q] = 0.0;
    bar[A_ph0][q] += bph;
    bar[A_ph_pert][q] += -bph - bpb;
    bar[A_ph_total][q] += bpb;
  }
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k);
    const double mref = mh_at(g, k, muts0), mcur = mh_at(g, k, mut);
    // th_work = mref (th0 - T0) - mcur (theta - T0)
    const double bt = bar[A_th_work][q];
    bar[A_th_work][q] = 0.0;
    bar[A_th0][q] += bt * mref;
    bar[A_theta][q] += -bt * mcur;
    b_muts0 += bt * (val[A_th0][q] - kT0) * g.c1h[k];
    b_mut += -bt * (val[A_theta][q] - kT0) * g.c1h[k];
    // c2a = (cpovcv (pb + p_pert)) / alt_ac; alt_ac = max(alt_raw, min); pb = p_total - p_pert
    const double pb = val[A_p_total][q] - val[A_p_pert][q];
    const bool alt_ok = val[A_alt_raw][q] > kMinAcousticAlt;
    const double alt_ac = alt_ok ? val[A_alt_raw][q] : kMinAcousticAlt;
    const double bc = bar[A_c2a][q];
    bar[A_c2a][q] = 0.0;
    const double b_sum = bc * kCpovcv / alt_ac;
    double b_alt = bar[A_alt][q] - bc * ((kCpovcv * (pb + val[A_p_pert][q])) / alt_ac) / alt_ac;
    bar[A_alt][q] = 0.0;
    const double b_pb = bar[A_pb][q] + b_sum;
    bar[A_pb][q] = 0.0;
    bar[A_p_pert][q] += b_sum - b_pb;
    bar[A_p_total][q] += b_pb;
    if (alt_ok) bar[A_alt_raw][q] += b_alt;
  }
  bar[A_mut][c] += b_mut;
  bar[A_muts0][c] += b_muts0;
}

// S3: u_work (ydir false) or v_work (ydir true) on face (j, i):
// work = (mref x0 - mcur x) / msf.
__host__ __device__ inline void adj_setup_s3_face(const DGrid& g, const View& V, bool ydir, int j, int i) {
  const int nz = g.nz;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const size_t f2 = ydir ? i2v(g, j, i) : i2u(g, j, i);
  const int wk = ydir ? A_v_work : A_u_work, x = ydir ? A_v : A_u, x0 = ydir ? A_v0 : A_u0;
  const int ms = ydir ? A_muvs : A_muus, mc = ydir ? A_muv : A_muu;
  const double msf = ydir ? g.msfvx[f2] : g.msfuy[f2];
  const double mus = val[ms][f2], muc = val[mc][f2];
  double b_ms = 0.0, b_mc = 0.0;
  for (int k = 0; k < nz; ++k) {
    const size_t q = ydir ? iv(g, k, j, i) : iu(g, k, j, i);
    const double b = bar[wk][q] / msf;
    bar[wk][q] = 0.0;
    bar[x0][q] += b * mh_at(g, k, mus);
    bar[x][q] += -b * mh_at(g, k, muc);
    b_ms += b * val[x0][q] * g.c1h[k];
    b_mc += -b * val[x][q] * g.c1h[k];
  }
  bar[ms][f2] += b_ms;
  bar[mc][f2] += b_mc;
}

// S4: the face averages (_u/_v_face_average_2d: edge faces copy the edge
// cell, inner faces average) of mut and muts0, gathered per cell, then the
// 2-D prep of column (j, i).  Face bars are cleared by adj_setup_s4_clear.
__host__ __device__ inline void adj_setup_s4_column(const DGrid& g, int rk_step, const View& V, int j, int i) {
  const int nx = g.nx, ny = g.ny;
  const size_t c = i2(g, j, i);
  double* const* bar = V.bar;
  auto gather_x = [&](int arr) {
    double b = 0.0;
    b += (i == 0) ? bar[arr][i2u(g, j, 0)] : 0.5 * bar[arr][i2u(g, j, i)];
    b += (i == nx - 1) ? bar[arr][i2u(g, j, nx)] : 0.5 * bar[arr][i2u(g, j, i + 1)];
    return b;
  };
  auto gather_y = [&](int arr) {
    double b = 0.0;
    b += (j == 0) ? bar[arr][i2v(g, 0, i)] : 0.5 * bar[arr][i2v(g, j, i)];
    b += (j == ny - 1) ? bar[arr][i2v(g, ny, i)] : 0.5 * bar[arr][i2v(g, j + 1, i)];
    return b;
  };
  const double b_mut = bar[A_mut][c] + gather_x(A_muu) + gather_y(A_muv);
  const double b_muts0 = bar[A_muts0][c] + gather_x(A_muus) + gather_y(A_muvs);
  // mub = mu_total - mu_pert; mu_work = rk1 ? 0 : mu0 - mu_pert; mut = mub + mu_pert;
  // muts0 = mut + mu_work; mu = mu_pert + mu_work
  const double b_mu = bar[A_mu][c];
  const double b_mw = bar[A_mu_work0][c] + b_muts0 + b_mu;
  const double b_mt = b_mut + b_muts0;
  const double b_mub = bar[A_mub][c] + b_mt;
  double b_mup = b_mt + b_mu - b_mub;
  if (rk_step != 1) {
    bar[A_mu0][c] += b_mw;
    b_mup -= b_mw;
  }
  bar[A_mu_total][c] += b_mub;
  bar[A_mu_pert][c] += b_mup;
  bar[A_mut][c] = 0.0;
  bar[A_muts0][c] = 0.0;
  bar[A_mu][c] = 0.0;
  bar[A_mu_work0][c] = 0.0;
  bar[A_mub][c] = 0.0;
}
__host__ __device__ inline void adj_setup_s4_clear(const DGrid& g, const View& V, bool ydir, int j, int i) {
  const size_t f2 = ydir ? i2v(g, j, i) : i2u(g, j, i);
  V.bar[ydir ? A_muv : A_muu][f2] = 0.0;
  V.bar[ydir ? A_muvs : A_muus][f2] = 0.0;
}

// S5: large-step PGF on u face (j, i) or v face; ut / vt are read-modify-write.
template <class Acc>
__host__ __device__ inline void adj_pgf_face(const DGrid& g, const View& V, bool ydir, int j, int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  int jl, il, jr, ir;
  size_t f2;
  double msf, rd;
  if (!ydir) {
    jl = jr = j;
    il = clamp_h(i - 1, 0, nx - 1);
    ir = clamp_h(i, 0, nx - 1);
    f2 = i2u(g, j, i);
    msf = g.msfux[f2] / g.msfuy[f2];
    rd = g.rdx;
  } else {
    il = ir = i;
    jl = clamp_h(j - 1, 0, ny - 1);
    jr = clamp_h(j, 0, ny - 1);
    f2 = i2v(g, j, i);
    msf = g.msfvy[f2] / g.msfvx[f2];
    rd = g.rdy;
  }
  const size_t cl = i2(g, jl, il), cr = i2(g, jr, ir);
  auto L = [=](int k) { return im(g, k, jl, il); };
  auto R = [=](int k) { return im(g, k, jr, ir); };
  const double* pp = val[A_p_pert];
  const double mub_l = val[A_mu_total][cl] - val[A_mu_pert][cl], mub_r = val[A_mu_total][cr] - val[A_mu_pert][cr];
  const double mface = 0.5 * ((mub_l + val[A_mu_pert][cl]) + (mub_r + val[A_mu_pert][cr]));
  auto ps = [=](int kk) { return pp[L(kk)] + pp[R(kk)]; };
  auto dpn = [=](int kk) {
    if (kk == 0) return 0.5 * ((g.cf1 * ps(0) + g.cf2 * ps(1)) + g.cf3 * ps(2));
    if (kk == nz) return 0.5 * (g.cfn * ps(nz - 1) + g.cfn1 * ps(nz - 2));
    return 0.5 * (g.fnm[kk] * ps(kk) + g.fnp[kk] * ps(kk - 1));
  };
  auto add_ps = [&](int kk, double v) {
    acc(&bar[A_p_pert][L(kk)], v);
    acc(&bar[A_p_pert][R(kk)], v);
  };
  auto dpn_bar = [&](int k, double b) {
    const double h = 0.5 * b;
    if (k == 0) {
      add_ps(0, h * g.cf1);
      add_ps(1, h * g.cf2);
      add_ps(2, h * g.cf3);
    } else if (k == nz) {
      add_ps(nz - 1, h * g.cfn);
      add_ps(nz - 2, h * g.cfn1);
    } else {
      add_ps(k, h * g.fnm[k]);
      add_ps(k - 1, h * g.fnp[k]);
    }
  };
  double b_mface = 0.0, b_musum = 0.0, carry = 0.0;
  for (int k = 0; k < nz; ++k) {
    const size_t q = ydir ? iv(g, k, j, i) : iu(g, k, j, i);
    const double b_dpx = -1.0 * bar[ydir ? A_vt : A_ut][q];
    const double mass = mh_at(g, k, mface);
    const double C = (msf * 0.5) * rd;
    const double ph_term = (val[A_ph_pert][R(k + 1)] - val[A_ph_pert][L(k + 1)]) +
                           (val[A_ph_pert][R(k)] - val[A_ph_pert][L(k)]);
    const double alt_s = val[A_alt_raw][L(k)] + val[A_alt_raw][R(k)];
    const double p_d = pp[R(k)] - pp[L(k)];
    const double pb_l = val[A_p_total][L(k)] - pp[L(k)], pb_r = val[A_p_total][R(k)] - pp[R(k)];
    const double al_s = val[A_al_ls][L(k)] + val[A_al_ls][R(k)];
    const double T = (ph_term + alt_s * p_d) + al_s * (pb_r - pb_l);
    const double Gk = (msf * rd) * (val[A_php_ls][R(k)] - val[A_php_ls][L(k)]);
    const double bracket = g.rdnw[k] * (dpn(k + 1) - dpn(k)) -
                           0.5 * (g.c1h[k] * (val[A_mu_pert][cl] + val[A_mu_pert][cr]));
    const double b_G = b_dpx * bracket, b_br = b_dpx * Gk;
    acc(&bar[A_php_ls][R(k)], b_G * (msf * rd));
    acc(&bar[A_php_ls][L(k)], -b_G * (msf * rd));
    dpn_bar(k, carry - b_br * g.rdnw[k]);
    carry = b_br * g.rdnw[k];
    b_musum += -b_br * 0.5 * g.c1h[k];
    b_mface += (b_dpx * C * T) * g.c1h[k];
    const double b_T = b_dpx * (C * mass);
    acc(&bar[A_ph_pert][R(k + 1)], b_T);
    acc(&bar[A_ph_pert][L(k + 1)], -b_T);
    acc(&bar[A_ph_pert][R(k)], b_T);
    acc(&bar[A_ph_pert][L(k)], -b_T);
    acc(&bar[A_alt_raw][L(k)], b_T * p_d);
    acc(&bar[A_alt_raw][R(k)], b_T * p_d);
    const double b_pd = b_T * alt_s;
    const double b_pbd = b_T * al_s;  // d/d(pb_r - pb_l)
    acc(&bar[A_al_ls][L(k)], b_T * (pb_r - pb_l));
    acc(&bar[A_al_ls][R(k)], b_T * (pb_r - pb_l));
    // p_d = pp_r - pp_l; pb = p_total - p_pert
    acc(&bar[A_p_pert][R(k)], b_pd - b_pbd);
    acc(&bar[A_p_pert][L(k)], -b_pd + b_pbd);
    acc(&bar[A_p_total][R(k)], b_pbd);
    acc(&bar[A_p_total][L(k)], -b_pbd);
  }
  dpn_bar(nz, carry);
  // mface = 0.5 ((mub_l + mu_pert_l) + (mub_r + mu_pert_r)), mub = mu_total - mu_pert
  const double bh = 0.5 * b_mface;
  acc(&bar[A_mu_total][cl], bh);
  acc(&bar[A_mu_total][cr], bh);
  acc(&bar[A_mu_pert][cl], b_musum);
  acc(&bar[A_mu_pert][cr], b_musum);
}

// S6: _absolute_diagnostics (al_ls, php_ls) and the EOS alt_raw on column (j, i).
template <class Acc>
__host__ __device__ inline void adj_setup_s6_column(const DGrid& g, const View& V, int j, int i, Acc acc) {
  const int nz = g.nz;
  const size_t c = i2(g, j, i);
  const double* const* val = V.val;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double mu_pert = val[A_mu_pert][c], mu_total = val[A_mu_total][c];
  const double mub = mu_total - mu_pert, muts = mu_total + mu_pert;
  double b_mub = 0.0, b_muts = 0.0, b_mup = 0.0;
  for (int k = 0; k < nz; ++k) {
    const size_t q = Q(k), q1 = Q(k + 1);
    const double phb_k = val[A_ph_total][q] - val[A_ph_pert][q];
    const double phb_k1 = val[A_ph_total][q1] - val[A_ph_pert][q1];
    const double mb = mh_at(g, k, mub);
    const bool b_ok = fabs(mb) > 1.0e-12;
    const double sb = b_ok ? mb : 1.0e-12;
    const double alb = ((-g.rdnw[k]) * (phb_k1 - phb_k)) / sb;
    const double mt = mh_at(g, k, muts);
    const bool t_ok = fabs(mt) > 1.0e-12;
    const double st = t_ok ? mt : 1.0e-12;
    const double mu_term = g.c1h[k] * mu_pert;
    const double A = alb * mu_term + g.rdnw[k] * (val[A_ph_pert][q1] - val[A_ph_pert][q]);
    // php = 0.5 (pht_k + pht_k1), pht = phb + ph_pert = ph_total (as written: (ph_total - ph_pert) + ph_pert)
    const double bp = 0.5 * bar[A_php_ls][q];
    bar[A_php_ls][q] = 0.0;
    double b_phb_k = bp, b_phb_k1 = bp;
    double b_php_k = bp, b_php_k1 = bp;
    // al = -A / st
    const double b_al = bar[A_al_ls][q];
    bar[A_al_ls][q] = 0.0;
    const double b_A = -b_al / st;
    if (t_ok) b_muts += (b_al * A / (st * st)) * g.c1h[k];
    const double b_alb = b_A * mu_term;
    b_mup += b_A * alb * g.c1h[k];
    b_php_k1 += b_A * g.rdnw[k];
    b_php_k += -b_A * g.rdnw[k];
    // alb = (-rdnw (phb_k1 - phb_k)) / sb
    const double b_d = b_alb * (-g.rdnw[k]) / sb;
    if (b_ok) b_mub += (-b_alb * alb / sb) * g.c1h[k];
    b_phb_k1 += b_d;
    b_phb_k += -b_d;
    acc(&bar[A_ph_total][q], b_phb_k);
    acc(&bar[A_ph_pert][q], b_php_k - b_phb_k);
    acc(&bar[A_ph_total][q1], b_phb_k1);
    acc(&bar[A_ph_pert][q1], b_php_k1 - b_phb_k1);
    // alt_raw = ((R/P0 theta) 1) pow(ps / P0, cvpm), ps = max(p_total, min)
    const double b_ar = bar[A_alt_raw][q];
    bar[A_alt_raw][q] = 0.0;
    const double pt = val[A_p_total][q];
    const bool p_ok = pt > kMinPressure;
    const double psv = p_ok ? pt : kMinPressure;
    const double x = psv / kP0;
    const double pw = wd_pow(x, kCvpm);
    const double th = val[A_theta][q];
    acc(&bar[A_theta][q], b_ar * (((kRd / kP0) * 1.0) * pw));
    if (p_ok) acc(&bar[A_p_total][q], b_ar * (((kRd / kP0) * th) * 1.0) * ((kCvpm * pw) / x) / kP0);
  }
  acc(&bar[A_mu_total][c], b_mub + b_muts);
  acc(&bar[A_mu_pert][c], -b_mub + b_muts + b_mup);
}

// ---------------------------------------------------------------------------
// S7: flux advection (flux_advection.py:53-88, 482-568, 1131-1482).
// flux5 / flux3 weights of their stencil values for a given upwind sign.
// ---------------------------------------------------------------------------
// flux5(q_im3, q_im2, q_im1, q_i, q_ip1, q_ip2, vel) = f6 - sgn(vel) corr
__host__ __device__ inline double flux5_val(const double* q, double s) {
  const double f6 = ((37.0 * (q[3] + q[2]) - 8.0 * (q[4] + q[1])) + (q[5] + q[0])) / 60.0;
  const double corr = (((q[5] - q[0]) - 5.0 * (q[4] - q[1])) + 10.0 * (q[3] - q[2])) / 60.0;
  return f6 - s * corr;
}
__host__ __device__ inline void flux5_w(double s, double* w) {
  w[0] = (1.0 + s) / 60.0;
  w[1] = (-8.0 - 5.0 * s) / 60.0;
  w[2] = (37.0 + 10.0 * s) / 60.0;
  w[3] = (37.0 - 10.0 * s) / 60.0;
  w[4] = (-8.0 + 5.0 * s) / 60.0;
  w[5] = (1.0 - s) / 60.0;
}
// flux3(q_km2, q_km1, q_k, q_kp1, velz) = f4 + sgn(velz) corr
__host__ __device__ inline double flux3_val(const double* q, double s) {
  const double f4 = (7.0 * (q[2] + q[1]) - (q[3] + q[0])) / 12.0;
  const double corr = ((q[3] - q[0]) - 3.0 * (q[2] - q[1])) / 12.0;
  return f4 + s * corr;
}
__host__ __device__ inline void flux3_w(double s, double* w) {
  w[0] = (-1.0 - s) / 12.0;
  w[1] = (7.0 + 3.0 * s) / 12.0;
  w[2] = (7.0 - 3.0 * s) / 12.0;
  w[3] = (-1.0 + s) / 12.0;
}

enum AdvKind : int { ADV_SCALAR = 0, ADV_U = 1, ADV_V = 2, ADV_W = 3 };

// Horizontal part at one output point (k, j, i) of kind's tendency: the bar
// of the tendency there goes to the advected field (theta - T0 goes to
// theta) and to ru / rv.  The tendency bar itself is cleared by the
// vertical pass.
template <class Acc>
__host__ __device__ inline void adj_advect_h_point(const DGrid& g, const View& V, int kind, int k, int j, int i,
                                                   Acc acc) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const double* ru = val[A_ru];
  const double* rv = val[A_rv];
  const int ip = wrap_h(i + 1, nx), imm = wrap_h(i - 1, nx), jp = wrap_h(j + 1, ny), jm = wrap_h(j - 1, ny);
  double bt;
  double msf;
  int fld;
  if (kind == ADV_SCALAR) {
    bt = bar[A_tht][im(g, k, j, i)];
    msf = g.msftx[i2(g, j, i)];
    fld = A_theta;
  } else if (kind == ADV_U) {
    bt = bar[A_ut][iu(g, k, j, i)];
    msf = g.msfux[i2u(g, j, i)];
    fld = A_u;
  } else if (kind == ADV_V) {
    bt = bar[A_vt][iv(g, k, j, i)];
    msf = g.msfvy[i2v(g, j, i)];
    fld = A_v;
  } else {
    bt = bar[A_wt][im(g, k, j, i)];
    msf = g.msftx[i2(g, j, i)];
    fld = A_w;
  }
  if (bt == 0.0) return;
  auto fidx = [=](int kk, int jj, int ii) -> size_t {
    if (kind == ADV_U) return iu(g, kk, jj, ii);
    if (kind == ADV_V) return iv(g, kk, jj, ii);
    return im(g, kk, jj, ii);
  };
  auto fval = [=](int jj, int ii) {
    const double x = val[fld][fidx(k, jj, ii)];
    return kind == ADV_SCALAR ? x - kT0 : x;
  };
  // mass_to_full for w levels
  auto m2f = [=](const double* r, int jj, int ii) {
    if (k == 0) return r[im(g, 0, jj, ii)];
    if (k == nz) return r[im(g, nz - 1, jj, ii)];
    return g.fnm[k] * r[im(g, k, jj, ii)] + g.fnp[k] * r[im(g, k - 1, jj, ii)];
  };
  auto m2f_bar = [&](int arr, int jj, int ii, double b) {
    if (k == 0) {
      acc(&bar[arr][im(g, 0, jj, ii)], b);
    } else if (k == nz) {
      acc(&bar[arr][im(g, nz - 1, jj, ii)], b);
    } else {
      acc(&bar[arr][im(g, k, jj, ii)], b * g.fnm[k]);
      acc(&bar[arr][im(g, k - 1, jj, ii)], b * g.fnp[k]);
    }
  };
  // one face flux vel * flux5(stencil) with its bar bf: stencil along x
  // (row j, cells i0-3 .. i0+2 periodic) or along y (column i)
  auto face = [&](bool ydir, int i0, double vel, double bf, double* b_vel) {
    double q[6];
    for (int t = 0; t < 6; ++t) q[t] = ydir ? fval(wrap_h(i0 + t - 3, ny), i) : fval(j, wrap_h(i0 + t - 3, nx));
    const double s = sgn0(vel);
    *b_vel += bf * flux5_val(q, s);
    double w[6];
    flux5_w(s, w);
    for (int t = 0; t < 6; ++t) {
      const size_t idx = ydir ? fidx(k, wrap_h(i0 + t - 3, ny), i) : fidx(k, j, wrap_h(i0 + t - 3, nx));
      acc(&bar[fld][idx], bf * vel * w[t]);
    }
  };
  // x direction
  {
    const double Kx = (-msf) * g.rdx;
    const double b1 = bt * Kx, b0 = -bt * Kx;
    double v0, v1, bv0 = 0.0, bv1 = 0.0;
    if (kind == ADV_SCALAR) {
      v0 = ru[im(g, k, j, i)];
      v1 = ru[im(g, k, j, ip)];
    } else if (kind == ADV_U) {
      v0 = 0.5 * (ru[im(g, k, j, i)] + ru[im(g, k, j, imm)]);
      v1 = 0.5 * (ru[im(g, k, j, ip)] + ru[im(g, k, j, i)]);
    } else if (kind == ADV_V) {
      v0 = 0.5 * (ru[im(g, k, j, i)] + ru[im(g, k, jm, i)]);
      v1 = 0.5 * (ru[im(g, k, j, ip)] + ru[im(g, k, jm, ip)]);
    } else {
      v0 = m2f(ru, j, i);
      v1 = m2f(ru, j, ip);
    }
    face(false, i, v0, b0, &bv0);
    face(false, ip, v1, b1, &bv1);
    if (kind == ADV_SCALAR) {
      acc(&bar[A_ru][im(g, k, j, i)], bv0);
      acc(&bar[A_ru][im(g, k, j, ip)], bv1);
    } else if (kind == ADV_U) {
      acc(&bar[A_ru][im(g, k, j, i)], 0.5 * (bv0 + bv1));
      acc(&bar[A_ru][im(g, k, j, imm)], 0.5 * bv0);
      acc(&bar[A_ru][im(g, k, j, ip)], 0.5 * bv1);
    } else if (kind == ADV_V) {
      acc(&bar[A_ru][im(g, k, j, i)], 0.5 * bv0);
      acc(&bar[A_ru][im(g, k, jm, i)], 0.5 * bv0);
      acc(&bar[A_ru][im(g, k, j, ip)], 0.5 * bv1);
      acc(&bar[A_ru][im(g, k, jm, ip)], 0.5 * bv1);
    } else {
      m2f_bar(A_ru, j, i, bv0);
      m2f_bar(A_ru, j, ip, bv1);
    }
  }
  // y direction (u and w skip it on a single row)
  if (kind == ADV_SCALAR || kind == ADV_V || ny > 1) {
    const double Ky = msf * g.rdy;
    const double b1 = -bt * Ky, b0 = bt * Ky;
    double v0, v1, bv0 = 0.0, bv1 = 0.0;
    if (kind == ADV_SCALAR) {
      v0 = rv[im(g, k, j, i)];
      v1 = rv[im(g, k, jp, i)];
    } else if (kind == ADV_U) {
      v0 = 0.5 * (rv[im(g, k, j, i)] + rv[im(g, k, j, imm)]);
      v1 = 0.5 * (rv[im(g, k, jp, i)] + rv[im(g, k, jp, imm)]);
    } else if (kind == ADV_V) {
      v0 = 0.5 * (rv[im(g, k, j, i)] + rv[im(g, k, jm, i)]);
      v1 = 0.5 * (rv[im(g, k, jp, i)] + rv[im(g, k, j, i)]);
    } else {
      v0 = m2f(rv, j, i);
      v1 = m2f(rv, jp, i);
    }
    face(true, j, v0, b0, &bv0);
    face(true, jp, v1, b1, &bv1);
    if (kind == ADV_SCALAR) {
      acc(&bar[A_rv][im(g, k, j, i)], bv0);
      acc(&bar[A_rv][im(g, k, jp, i)], bv1);
    } else if (kind == ADV_U) {
      acc(&bar[A_rv][im(g, k, j, i)], 0.5 * bv0);
      acc(&bar[A_rv][im(g, k, j, imm)], 0.5 * bv0);
      acc(&bar[A_rv][im(g, k, jp, i)], 0.5 * bv1);
      acc(&bar[A_rv][im(g, k, jp, imm)], 0.5 * bv1);
    } else if (kind == ADV_V) {
      acc(&bar[A_rv][im(g, k, j, i)], 0.5 * (bv0 + bv1));
      acc(&bar[A_rv][im(g, k, jm, i)], 0.5 * bv0);
      acc(&bar[A_rv][im(g, k, jp, i)], 0.5 * bv1);
    } else {
      m2f_bar(A_rv, j, i, bv0);
      m2f_bar(A_rv, jp, i, bv1);
    }
  }
}

// Vertical part on column (j, i) of kind's tendency (u: face column i < nx,
// v: face row j < ny), then the tendency's bar is cleared.  The periodic
// copy of face nx (u) / row ny (v) from face 0 / row 0 is undone first by
// adj_advect_copy.
template <class Acc>
__host__ __device__ inline void adj_advect_v_column(const DGrid& g, const View& V, int kind, int j, int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const double* rom = val[A_rom];
  if (kind == ADV_W) {
    // vflux[k] for k = 1 .. nz-1 (vflux[0] = vflux[nz] = 0); wt[k] += -rdn[k] (vflux[k+1] - vflux[k])
    auto VF = [=](int k) { return 0.5 * (rom[im(g, k, j, i)] + rom[im(g, k - 1, j, i)]); };
    auto WW = [=](int k) { return val[A_w][im(g, k, j, i)]; };
    double b_prev = 0.0;  // bar of vflux[k+1] from level k
    for (int k = nz - 1; k >= 1; --k) {
      const double bt = bar[A_wt][im(g, k, j, i)];
      // vflux[k+1] gets -rdn[k] bt (added to the carried bar), vflux[k] gets +rdn[k] bt
      const double b_up = b_prev + (-g.rdn[k]) * bt;
      const int kf = k + 1;
      if (kf <= nz - 1) {
        // vflux[kf] (kf >= 2)
        const double vel = VF(kf);
        double q[4] = {WW(kf - 2), WW(kf - 1), WW(kf), WW(kf + 1)};
        const double s = sgn0(-vel);
        const double b_vel = b_up * flux3_val(q, s);
        double w[4];
        flux3_w(s, w);
        for (int t = 0; t < 4; ++t) acc(&bar[A_w][im(g, kf - 2 + t, j, i)], b_up * vel * w[t]);
        acc(&bar[A_rom][im(g, kf, j, i)], 0.5 * b_vel);
        acc(&bar[A_rom][im(g, kf - 1, j, i)], 0.5 * b_vel);
      }
      b_prev = g.rdn[k] * bt;
    }
    // vflux[1] = (VF(1) 0.5) (w[1] + w[0])
    {
      const double b1 = b_prev;
      const double vel = VF(1);
      acc(&bar[A_w][im(g, 1, j, i)], b1 * vel * 0.5);
      acc(&bar[A_w][im(g, 0, j, i)], b1 * vel * 0.5);
      const double b_vel = b1 * 0.5 * (WW(1) + WW(0));
      acc(&bar[A_rom][im(g, 1, j, i)], 0.5 * b_vel);
      acc(&bar[A_rom][im(g, 0, j, i)], 0.5 * b_vel);
    }
    for (int k = 0; k <= nz; ++k) bar[A_wt][im(g, k, j, i)] = 0.0;
    return;
  }
  // mass-level kinds: q column and transport
  auto qv = [=](int k) {
    if (kind == ADV_SCALAR) return val[A_theta][im(g, k, j, i)] - kT0;
    if (kind == ADV_U) return val[A_u][iu(g, k, j, i)];
    return val[A_v][iv(g, k, j, i)];
  };
  auto qidx = [=](int k) -> size_t {
    if (kind == ADV_U) return iu(g, k, j, i);
    if (kind == ADV_V) return iv(g, k, j, i);
    return im(g, k, j, i);
  };
  const int fld = kind == ADV_SCALAR ? A_theta : (kind == ADV_U ? A_u : A_v);
  const int tarr = kind == ADV_SCALAR ? A_tht : (kind == ADV_U ? A_ut : A_vt);
  // transport on the column's faces and where its bar goes
  const int jo = kind == ADV_V ? wrap_h(j - 1, ny) : j, io = kind == ADV_U ? wrap_h(i - 1, nx) : i;
  auto romq = [=](int k) {
    if (kind == ADV_SCALAR) return rom[im(g, k, j, i)];
    return 0.5 * (rom[im(g, k, j, i)] + rom[im(g, k, jo, io)]);
  };
  auto romq_bar = [&](int k, double b) {
    if (kind == ADV_SCALAR) {
      acc(&bar[A_rom][im(g, k, j, i)], b);
    } else {
      acc(&bar[A_rom][im(g, k, j, i)], 0.5 * b);
      acc(&bar[A_rom][im(g, k, jo, io)], 0.5 * b);
    }
  };
  const double ratio = kind == ADV_V ? g.msfvy[i2v(g, j, i)] / g.msfvx[i2v(g, j, i)] : 1.0;
  // tend[k] += ratio (-rdnw[k]) (vf[k+1] - vf[k]); the bar of vf[kf] is
  // ratio rdnw[kf] bt[kf] - ratio rdnw[kf-1] bt[kf-1]
  for (int kf = 1; kf <= nz - 1; ++kf) {
    const double b_vf = ratio * ((-g.rdnw[kf - 1]) * bar[tarr][qidx(kf - 1)] + g.rdnw[kf] * bar[tarr][qidx(kf)]);
    if (kf == 1 || kf == nz - 1) {
      const double lin = g.fnm[kf] * qv(kf) + g.fnp[kf] * qv(kf - 1);
      const double r = romq(kf);
      romq_bar(kf, b_vf * lin);
      acc(&bar[fld][qidx(kf)], b_vf * r * g.fnm[kf]);
      acc(&bar[fld][qidx(kf - 1)], b_vf * r * g.fnp[kf]);
    } else {
      const double r = romq(kf);
      double q[4] = {qv(kf - 2), qv(kf - 1), qv(kf), qv(kf + 1)};
      const double s = sgn0(-r);
      romq_bar(kf, b_vf * flux3_val(q, s));
      double w[4];
      flux3_w(s, w);
      for (int t = 0; t < 4; ++t) acc(&bar[fld][qidx(kf - 2 + t)], b_vf * r * w[t]);
    }
  }
  for (int k = 0; k < nz; ++k) bar[tarr][qidx(k)] = 0.0;
}

// ut[k, j, nx] = ut[k, j, 0] (u) and vt[k, ny, i] = vt[k, 0, i] (v), undone
// before the horizontal pass of u / v.
__host__ __device__ inline void adj_advect_copy(const DGrid& g, const View& V, bool ydir, int k, int r) {
  if (!ydir) {
    V.bar[A_ut][iu(g, k, r, 0)] += V.bar[A_ut][iu(g, k, r, g.nx)];
    V.bar[A_ut][iu(g, k, r, g.nx)] = 0.0;
  } else {
    V.bar[A_vt][iv(g, k, 0, r)] += V.bar[A_vt][iv(g, k, g.ny, r)];
    V.bar[A_vt][iv(g, k, g.ny, r)] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// S8 / S9: couple_velocities_periodic (flux_advection.py:394-470).
// ---------------------------------------------------------------------------
template <class Acc>
__host__ __device__ inline void adj_couple_column(const DGrid& g, const View& V, int j, int i, Acc acc) {
  const int nz = g.nz, nx = g.nx, ny = g.ny;
  double* const* bar = V.bar;
  auto Q = [=](int k) { return im(g, k, j, i); };
  const double msftx = g.msftx[i2(g, j, i)];
  // rom[k+1] = cum_k, cum_k = sum_{k' <= k} incr_k', incr = -((dnw c1h) dmdt) - divv
  double b_cum = 0.0, b_dmdt = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    if (k + 1 < nz) b_cum += bar[A_rom][Q(k + 1)];
    b_dmdt += -b_cum * (g.dnw[k] * g.c1h[k]);
    bar[A_divv][Q(k)] += -b_cum;
  }
  for (int k = 0; k <= nz; ++k) bar[A_rom][Q(k)] = 0.0;
  const int ip = wrap_h(i + 1, nx), jp = wrap_h(j + 1, ny);
  for (int k = 0; k < nz; ++k) {
    const double b_d = bar[A_divv][Q(k)] + b_dmdt;
    bar[A_divv][Q(k)] = 0.0;
    const double b_in = b_d * (g.dnw[k] * msftx);
    acc(&bar[A_ru][im(g, k, j, ip)], b_in * g.rdx);
    acc(&bar[A_ru][Q(k)], -b_in * g.rdx);
    acc(&bar[A_rv][im(g, k, jp, i)], b_in * g.rdy);
    acc(&bar[A_rv][Q(k)], -b_in * g.rdy);
  }
}

// ru = (mass_u u) / msfuy, mass_u from the x average of mu_total; rv likewise in y.
template <class Acc>
__host__ __device__ inline void adj_couple_point(const DGrid& g, const View& V, int k, int j, int i, Acc acc) {
  const int nx = g.nx, ny = g.ny;
  const double* const* val = V.val;
  double* const* bar = V.bar;
  const double* mu = val[A_mu_total];
  const size_t q = im(g, k, j, i);
  const int imm = wrap_h(i - 1, nx), jm = wrap_h(j - 1, ny);
  const double bru = bar[A_ru][q], brv = bar[A_rv][q];
  bar[A_ru][q] = 0.0;
  bar[A_rv][q] = 0.0;
  const double muu = 0.5 * (mu[i2(g, j, i)] + mu[i2(g, j, imm)]);
  const double muv = 0.5 * (mu[i2(g, j, i)] + mu[i2(g, jm, i)]);
  const double msu = g.msfuy[i2u(g, j, i)], msv = g.msfvx[i2v(g, j, i)];
  const double uu = val[A_u][iu(g, k, j, i)], vv = val[A_v][iv(g, k, j, i)];
  acc(&bar[A_u][iu(g, k, j, i)], bru * mh_at(g, k, muu) / msu);
  acc(&bar[A_v][iv(g, k, j, i)], brv * mh_at(g, k, muv) / msv);
  const double bmu_u = bru * uu / msu * g.c1h[k] * 0.5, bmu_v = brv * vv / msv * g.c1h[k] * 0.5;
  acc(&bar[A_mu_total][i2(g, j, i)], bmu_u + bmu_v);
  acc(&bar[A_mu_total][i2(g, j, imm)], bmu_u);
  acc(&bar[A_mu_total][i2(g, jm, i)], bmu_v);
}

}  // namespace adj
}  // namespace wd
