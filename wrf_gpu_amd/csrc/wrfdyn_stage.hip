// RK-stage setup kernels (once per RK stage): transport velocities, flux-form
// advection, large-step PGF, small_step_prep, calc_p_rho step 0, the stage
// rw_tend / ph_tend and the calc_coef_w tridiagonal coefficients.
//
// Every kernel computes each output value with the reference's expression
// order (see the file:line cited per kernel); build with -ffp-contract=off.
#include "wrfdyn_kernels.h"

namespace wd {

namespace {

__device__ inline void decompose(size_t idx, int nxx, int nyy, int& k, int& j, int& i) {
  i = (int)(idx % (size_t)nxx);
  size_t t = idx / (size_t)nxx;
  j = (int)(t % (size_t)nyy);
  k = (int)(t / (size_t)nyy);
}

// ---------------------------------------------------------------------------
// couple_velocities_periodic ru/rv, dynamics/flux_advection.py:394-428
// ---------------------------------------------------------------------------
__global__ void k_ruv(DGrid g, const double* __restrict__ u, const double* __restrict__ v,
                      const double* __restrict__ mu, double* __restrict__ ru, double* __restrict__ rv) {
  size_t n = (size_t)g.nz * g.ny * g.nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, g.nx, g.ny, k, j, i);
  double muu = 0.5 * (mu[i2(g, j, i)] + mu[i2(g, j, wrapi(i - 1, g.nx))]);
  double muv = 0.5 * (mu[i2(g, j, i)] + mu[i2(g, wrapi(j - 1, g.ny), i)]);
  double mass_u = g.c1h[k] * muu + g.c2h[k];
  double mass_v = g.c1h[k] * muv + g.c2h[k];
  ru[idx] = mass_u * u[iu(g, k, j, i)] / g.msfuy[i2u(g, j, i)];
  rv[idx] = mass_v * v[iv(g, k, j, i)] / g.msfvx[i2v(g, j, i)];
}

// calc_ww_cp omega, dynamics/flux_advection.py:430-458 (one thread per column)
__global__ void k_rom(DGrid g, const double* __restrict__ ru, const double* __restrict__ rv,
                      double* __restrict__ rom) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= g.nx) return;
  int ip = wrapi(i + 1, g.nx), jp = wrapi(j + 1, g.ny);
  double msftx = g.msftx[i2(g, j, i)];
  double dmdt = 0.0;
  for (int k = 0; k < g.nz; ++k) {
    double d = (g.dnw[k] * msftx) * (g.rdx * (ru[im(g, k, j, ip)] - ru[im(g, k, j, i)]) +
                                     g.rdy * (rv[im(g, k, jp, i)] - rv[im(g, k, j, i)]));
    dmdt += d;
  }
  double cum = 0.0;
  rom[im(g, 0, j, i)] = 0.0;
  for (int k = 0; k < g.nz; ++k) {
    double d = (g.dnw[k] * msftx) * (g.rdx * (ru[im(g, k, j, ip)] - ru[im(g, k, j, i)]) +
                                     g.rdy * (rv[im(g, k, jp, i)] - rv[im(g, k, j, i)]));
    double incr = -((g.dnw[k] * g.c1h[k]) * dmdt) - d;
    cum = (k == 0) ? incr : cum + incr;
    if (k + 1 < g.nz) rom[im(g, k + 1, j, i)] = cum;
  }
  rom[im(g, g.nz, j, i)] = 0.0;
}

// ---------------------------------------------------------------------------
// Stage-entry EOS + _absolute_diagnostics (core/rk_addtend_dry.py:127-229,
// hypsometric_opt=1): alt (shared with small_step_prep's base-less
// diagnose_pressure_al_alt, acoustic_wrf.py:354-358), large-step al and php.
// ---------------------------------------------------------------------------
__global__ void k_diag(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.nz * g.ny * g.nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, g.nx, g.ny, k, j, i);
  w.alt_raw[idx] = alt_from_theta_p(s.theta[idx], s.p_total[idx]);
  size_t c = i2(g, j, i);
  double mu_pert = s.mu_pert[c];
  double mu_total = s.mu_total[c];
  double mub = mu_total - mu_pert;
  double muts = mu_total + mu_pert;
  size_t q = idx, q1 = im(g, k + 1, j, i);
  double phb_k = s.ph_total[q] - s.ph_pert[q];
  double phb_k1 = s.ph_total[q1] - s.ph_pert[q1];
  double mass_base = g.c1h[k] * mub + g.c2h[k];
  double safe_base = safe_pos(mass_base, 1.0e-12);
  double alb = ((-g.rdnw[k]) * (phb_k1 - phb_k)) / safe_base;
  double mass_t = g.c1h[k] * muts + g.c2h[k];
  double safe_t = safe_pos(mass_t, 1.0e-12);
  double mu_term = g.c1h[k] * mu_pert;
  w.al_ls[q] = -(alb * mu_term + g.rdnw[k] * (s.ph_pert[q1] - s.ph_pert[q])) / safe_t;
  double pht_k = phb_k + s.ph_pert[q];
  double pht_k1 = phb_k1 + s.ph_pert[q1];
  w.php_ls[q] = 0.5 * (pht_k + pht_k1);
}

// dpn face value from pair sums of a mass-level pressure p at columns (a, b)
// (large-step PGF form, rk_addtend_dry.py:314-327, top_lid with cfn/cfn1).
__device__ inline double dpn_ls(const DGrid& g, const double* __restrict__ p, int kk, size_t ca,
                                size_t cb, size_t plane) {
#define PS(k) (p[(size_t)(k) * plane + ca] + p[(size_t)(k) * plane + cb])
  if (kk == 0) return 0.5 * ((g.cf1 * PS(0) + g.cf2 * PS(1)) + g.cf3 * PS(2));
  if (kk == g.nz) return 0.5 * (g.cfn * PS(g.nz - 1) + g.cfn1 * PS(g.nz - 2));
  return 0.5 * (g.fnm[kk] * PS(kk) + g.fnp[kk] * PS(kk - 1));
#undef PS
}

// _vertical_flux_div_3 face flux for one column (flux_advection.py:1362-1407);
// q(k) and romq(k) are accessed through the functors.
template <class Q, class R>
__device__ inline double vflux3_face(const DGrid& g, int kf, Q q, R romq) {
  const int nz = g.nz;
  if (kf == 0 || kf == nz) return 0.0;
  if (kf == 1) return romq(1) * (g.fnm[1] * q(1) + g.fnp[1] * q(0));
  if (kf == nz - 1) return romq(nz - 1) * (g.fnm[nz - 1] * q(nz - 1) + g.fnp[nz - 1] * q(nz - 2));
  double r = romq(kf);
  return r * flux3(q(kf - 2), q(kf - 1), q(kf), q(kf + 1), -r);
}

// ---------------------------------------------------------------------------
// advect_u_flux (flux_advection.py:1131-1202) + large_step_horizontal_pgf x
// (rk_addtend_dry.py:329-343, 357-361) on u faces (nz, ny, nx+1).
// ---------------------------------------------------------------------------
__global__ void k_tend_u(DGrid g, DState s, DWork w) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)nz * ny * (nx + 1);
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, nx + 1, ny, k, j, i);
  const double* u = s.u;
  // --- advection on the collapsed periodic face ia ---
  int ia = (i == nx) ? 0 : i;
  int ip = wrapi(ia + 1, nx), imn = wrapi(ia - 1, nx);
  auto UX = [&](int sft) { return u[iu(g, k, j, wrapi(ia + sft, nx))]; };
  double velx0 = 0.5 * (w.ru[im(g, k, j, ia)] + w.ru[im(g, k, j, imn)]);
  double velx1 = 0.5 * (w.ru[im(g, k, j, ip)] + w.ru[im(g, k, j, ia)]);
  double fq0 = velx0 * flux5(UX(-3), UX(-2), UX(-1), UX(0), UX(1), UX(2), velx0);
  double fq1 = velx1 * flux5(UX(-2), UX(-1), UX(0), UX(1), UX(2), UX(3), velx1);
  double msf = g.msfux[i2u(g, j, ia)];
  double t = ((-msf) * g.rdx) * (fq1 - fq0);
  if (ny > 1) {
    int jp = wrapi(j + 1, ny);
    auto UY = [&](int jj, int sft) { return u[iu(g, k, wrapi(jj + sft, ny), ia)]; };
    double vy0 = 0.5 * (w.rv[im(g, k, j, ia)] + w.rv[im(g, k, j, imn)]);
    double vy1 = 0.5 * (w.rv[im(g, k, jp, ia)] + w.rv[im(g, k, jp, imn)]);
    double g0 = vy0 * flux5(UY(j, -3), UY(j, -2), UY(j, -1), UY(j, 0), UY(j, 1), UY(j, 2), vy0);
    double g1 = vy1 * flux5(UY(jp, -3), UY(jp, -2), UY(jp, -1), UY(jp, 0), UY(jp, 1), UY(jp, 2), vy1);
    t = t - (msf * g.rdy) * (g1 - g0);
  }
  {
    auto q = [&](int kk) { return u[iu(g, kk, j, ia)]; };
    auto romq = [&](int kk) { return 0.5 * (w.rom[im(g, kk, j, ia)] + w.rom[im(g, kk, j, imn)]); };
    double vlo = vflux3_face(g, k, q, romq);
    double vhi = vflux3_face(g, k + 1, q, romq);
    t = t + (-g.rdnw[k]) * (vhi - vlo);
  }
  // --- large-step horizontal PGF at face i (edge-padded pairs) ---
  int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
  size_t cl = i2(g, j, il), cr = i2(g, j, ir);
  size_t plane = (size_t)ny * nx;
  double msf_u = g.msfux[i2u(g, j, i)] / g.msfuy[i2u(g, j, i)];
  double mub_l = s.mu_total[cl] - s.mu_pert[cl];
  double mub_r = s.mu_total[cr] - s.mu_pert[cr];
  double mt_l = mub_l + s.mu_pert[cl];
  double mt_r = mub_r + s.mu_pert[cr];
  double muu = 0.5 * (mt_l + mt_r);
  double mu_l = s.mu_pert[cl], mu_r = s.mu_pert[cr];
  double dpn_k = dpn_ls(g, s.p_pert, k, cl, cr, plane);
  double dpn_k1 = dpn_ls(g, s.p_pert, k + 1, cl, cr, plane);
  size_t ql = im(g, k, j, il), qr = im(g, k, j, ir);
  size_t ql1 = im(g, k + 1, j, il), qr1 = im(g, k + 1, j, ir);
  double mass_u = g.c1h[k] * muu + g.c2h[k];
  double ph_term = (s.ph_pert[qr1] - s.ph_pert[ql1]) + (s.ph_pert[qr] - s.ph_pert[ql]);
  double p_term = (w.alt_raw[ql] + w.alt_raw[qr]) * (s.p_pert[qr] - s.p_pert[ql]);
  double pb_l = s.p_total[ql] - s.p_pert[ql];
  double pb_r = s.p_total[qr] - s.p_pert[qr];
  double pb_term = (w.al_ls[ql] + w.al_ls[qr]) * (pb_r - pb_l);
  double dpx = (((msf_u * 0.5) * g.rdx) * mass_u) * ((ph_term + p_term) + pb_term);
  double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (mu_l + mu_r));
  dpx = dpx + ((msf_u * g.rdx) * (w.php_ls[qr] - w.php_ls[ql])) * bracket;
  w.ut[idx] = t + (-1.0 * dpx);
}

// advect_v_flux (flux_advection.py:1205-1274) + large-step PGF y
// (rk_addtend_dry.py:345-361) on v faces (nz, ny+1, nx).
__global__ void k_tend_v(DGrid g, DState s, DWork w) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)nz * (ny + 1) * nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, nx, ny + 1, k, j, i);
  const double* v = s.v;
  int ja = (j == ny) ? 0 : j;
  int jm = wrapi(ja - 1, ny), jp = wrapi(ja + 1, ny), ip = wrapi(i + 1, nx);
  auto VX = [&](int sft) { return v[iv(g, k, ja, wrapi(i + sft, nx))]; };
  auto VY = [&](int jj, int sft) { return v[iv(g, k, wrapi(jj + sft, ny), i)]; };
  double vx0 = 0.5 * (w.ru[im(g, k, ja, i)] + w.ru[im(g, k, jm, i)]);
  double vx1 = 0.5 * (w.ru[im(g, k, ja, ip)] + w.ru[im(g, k, jm, ip)]);
  double fq0 = vx0 * flux5(VX(-3), VX(-2), VX(-1), VX(0), VX(1), VX(2), vx0);
  double fq1 = vx1 * flux5(VX(-2), VX(-1), VX(0), VX(1), VX(2), VX(3), vx1);
  double msfvy = g.msfvy[i2v(g, ja, i)];
  double t = ((-msfvy) * g.rdx) * (fq1 - fq0);
  double vy0 = 0.5 * (w.rv[im(g, k, ja, i)] + w.rv[im(g, k, jm, i)]);
  double vy1 = 0.5 * (w.rv[im(g, k, jp, i)] + w.rv[im(g, k, ja, i)]);
  double g0 = vy0 * flux5(VY(ja, -3), VY(ja, -2), VY(ja, -1), VY(ja, 0), VY(ja, 1), VY(ja, 2), vy0);
  double g1 = vy1 * flux5(VY(jp, -3), VY(jp, -2), VY(jp, -1), VY(jp, 0), VY(jp, 1), VY(jp, 2), vy1);
  t = t - (msfvy * g.rdy) * (g1 - g0);
  {
    auto q = [&](int kk) { return v[iv(g, kk, ja, i)]; };
    auto romq = [&](int kk) { return 0.5 * (w.rom[im(g, kk, ja, i)] + w.rom[im(g, kk, jm, i)]); };
    double ratio = g.msfvy[i2v(g, ja, i)] / g.msfvx[i2v(g, ja, i)];
    double vlo = vflux3_face(g, k, q, romq);
    double vhi = vflux3_face(g, k + 1, q, romq);
    double vd = (-g.rdnw[k]) * (vhi - vlo);
    t = t + ratio * vd;
  }
  // --- large-step PGF at v face j ---
  int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
  size_t cs = i2(g, js, i), cn = i2(g, jn, i);
  size_t plane = (size_t)ny * nx;
  double msf_v = g.msfvy[i2v(g, j, i)] / g.msfvx[i2v(g, j, i)];
  double mub_s = s.mu_total[cs] - s.mu_pert[cs];
  double mub_n = s.mu_total[cn] - s.mu_pert[cn];
  double mt_s = mub_s + s.mu_pert[cs];
  double mt_n = mub_n + s.mu_pert[cn];
  double muv = 0.5 * (mt_s + mt_n);
  double mu_s = s.mu_pert[cs], mu_n = s.mu_pert[cn];
  double dpn_k = dpn_ls(g, s.p_pert, k, cs, cn, plane);
  double dpn_k1 = dpn_ls(g, s.p_pert, k + 1, cs, cn, plane);
  size_t qs = im(g, k, js, i), qn = im(g, k, jn, i);
  size_t qs1 = im(g, k + 1, js, i), qn1 = im(g, k + 1, jn, i);
  double mass_v = g.c1h[k] * muv + g.c2h[k];
  double ph_term = (s.ph_pert[qn1] - s.ph_pert[qs1]) + (s.ph_pert[qn] - s.ph_pert[qs]);
  double p_term = (w.alt_raw[qs] + w.alt_raw[qn]) * (s.p_pert[qn] - s.p_pert[qs]);
  double pb_s = s.p_total[qs] - s.p_pert[qs];
  double pb_n = s.p_total[qn] - s.p_pert[qn];
  double pb_term = (w.al_ls[qs] + w.al_ls[qn]) * (pb_n - pb_s);
  double dpy = (((msf_v * 0.5) * g.rdy) * mass_v) * ((ph_term + p_term) + pb_term);
  double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (mu_s + mu_n));
  dpy = dpy + ((msf_v * g.rdy) * (w.php_ls[qn] - w.php_ls[qs])) * bracket;
  w.vt[idx] = t + (-1.0 * dpy);
}

// _mass_to_full_levels (flux_advection.py:1332-1348)
__device__ inline double mass_to_full(const DGrid& g, const double* __restrict__ r, int kf, int j, int i) {
  if (kf == 0) return r[im(g, 0, j, i)];
  if (kf == g.nz) return r[im(g, g.nz - 1, j, i)];
  return g.fnm[kf] * r[im(g, kf, j, i)] + g.fnp[kf] * r[im(g, kf - 1, j, i)];
}

// advect_w_flux + _vertical_flux_div_w with a rigid lid
// (flux_advection.py:1277-1329, 1410-1482) on w levels (nz+1, ny, nx).
__global__ void k_tend_w(DGrid g, DState s, DWork w) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)(nz + 1) * ny * nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, nx, ny, k, j, i);
  const double* ww = s.w;
  int ip = wrapi(i + 1, nx);
  auto WX = [&](int sft) { return ww[im(g, k, j, wrapi(i + sft, nx))]; };
  double vx0 = mass_to_full(g, w.ru, k, j, i), vx1 = mass_to_full(g, w.ru, k, j, ip);
  double fq0 = vx0 * flux5(WX(-3), WX(-2), WX(-1), WX(0), WX(1), WX(2), vx0);
  double fq1 = vx1 * flux5(WX(-2), WX(-1), WX(0), WX(1), WX(2), WX(3), vx1);
  double msf = g.msftx[i2(g, j, i)];
  double t = ((-msf) * g.rdx) * (fq1 - fq0);
  if (ny > 1) {
    int jp = wrapi(j + 1, ny);
    auto WY = [&](int jj, int sft) { return ww[im(g, k, wrapi(jj + sft, ny), i)]; };
    double vy0 = mass_to_full(g, w.rv, k, j, i), vy1 = mass_to_full(g, w.rv, k, jp, i);
    double g0 = vy0 * flux5(WY(j, -3), WY(j, -2), WY(j, -1), WY(j, 0), WY(j, 1), WY(j, 2), vy0);
    double g1 = vy1 * flux5(WY(jp, -3), WY(jp, -2), WY(jp, -1), WY(jp, 0), WY(jp, 1), WY(jp, 2), vy1);
    t = t - (msf * g.rdy) * (g1 - g0);
  }
  if (k >= 1 && k < nz) {
    auto VF = [&](int kk) { return 0.5 * (w.rom[im(g, kk, j, i)] + w.rom[im(g, kk - 1, j, i)]); };
    auto W = [&](int kk) { return ww[im(g, kk, j, i)]; };
    auto vflux = [&](int kk) -> double {
      if (kk <= 0 || kk >= nz) return 0.0;
      if (kk == 1) return (VF(1) * 0.5) * (W(1) + W(0));
      double vel = VF(kk);
      return vel * flux3(W(kk - 2), W(kk - 1), W(kk), W(kk + 1), -vel);
    };
    t = t + (-g.rdn[k]) * (vflux(k + 1) - vflux(k));
  } else {
    t = t + 0.0;
  }
  w.wt[idx] = t;
}

// advect_scalar_flux of theta-300 (flux_advection.py:482-568) on mass levels.
__global__ void k_tend_t(DGrid g, DState s, DWork w) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)nz * ny * nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, nx, ny, k, j, i);
  const double* th = s.theta;
  auto QX = [&](int sft) { return th[im(g, k, j, wrapi(i + sft, nx))] - kT0; };
  auto QY = [&](int jj, int sft) { return th[im(g, k, wrapi(jj + sft, ny), i)] - kT0; };
  int ip = wrapi(i + 1, nx), jp = wrapi(j + 1, ny);
  double vel0 = w.ru[im(g, k, j, i)], vel1 = w.ru[im(g, k, j, ip)];
  double fq0 = vel0 * flux5(QX(-3), QX(-2), QX(-1), QX(0), QX(1), QX(2), vel0);
  double fq1 = vel1 * flux5(QX(-2), QX(-1), QX(0), QX(1), QX(2), QX(3), vel1);
  double msftx = g.msftx[i2(g, j, i)];
  double t = ((-msftx) * g.rdx) * (fq1 - fq0);
  double vy0 = w.rv[im(g, k, j, i)], vy1 = w.rv[im(g, k, jp, i)];
  double gy0 = vy0 * flux5(QY(j, -3), QY(j, -2), QY(j, -1), QY(j, 0), QY(j, 1), QY(j, 2), vy0);
  double gy1 = vy1 * flux5(QY(jp, -3), QY(jp, -2), QY(jp, -1), QY(jp, 0), QY(jp, 1), QY(jp, 2), vy1);
  t = t - (msftx * g.rdy) * (gy1 - gy0);
  auto q = [&](int kk) { return th[im(g, kk, j, i)] - kT0; };
  auto vf = [&](int kf) -> double {
    if (kf == nz || kf == 0) return 0.0;
    if (kf == 1) return w.rom[im(g, 1, j, i)] * (g.fnm[1] * q(1) + g.fnp[1] * q(0));
    if (kf == nz - 1) return w.rom[im(g, nz - 1, j, i)] * (g.fnm[nz - 1] * q(nz - 1) + g.fnp[nz - 1] * q(nz - 2));
    double r = w.rom[im(g, kf, j, i)];
    return r * flux3(q(kf - 2), q(kf - 1), q(kf), q(kf + 1), -r);
  };
  t = t - g.rdnw[k] * (vf(k + 1) - vf(k));
  w.tht[idx] = t;
}

// ---------------------------------------------------------------------------
// small_step_prep_wrf, core/small_step_prep.py:216-394
// ---------------------------------------------------------------------------
__global__ void k_prep_2d(DGrid g, DState s, DWork w, int rk_step) {
  size_t n = (size_t)g.ny * g.nx;
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double mub = s.mu_total[c] - s.mu_pert[c];
  double mu_cur = s.mu_pert[c];
  double mu_ref = w.mu0[c];
  double mu_work = rk_step == 1 ? 0.0 : mu_ref - mu_cur;
  double mut = mub + mu_cur;
  double muts = mut + mu_work;
  w.mub[c] = mub;
  w.mut[c] = mut;
  w.mu_work0[c] = mu_work;
  w.muts0[c] = muts;
  w.muts[c] = muts;
  w.mu[c] = mu_cur + mu_work;
  w.muave[c] = 0.0;
}

// _u_face_average_2d / _v_face_average_2d (small_step_prep.py:77-88)
__global__ void k_prep_faces(DGrid g, DWork w) {
  const int nx = g.nx, ny = g.ny;
  size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < nu) {
    int i = (int)(idx % (size_t)(nx + 1)), j = (int)(idx / (size_t)(nx + 1));
    double a, b;
    if (i == 0) {
      a = w.mut[i2(g, j, 0)];
      b = w.muts0[i2(g, j, 0)];
    } else if (i == nx) {
      a = w.mut[i2(g, j, nx - 1)];
      b = w.muts0[i2(g, j, nx - 1)];
    } else {
      a = 0.5 * (w.mut[i2(g, j, i - 1)] + w.mut[i2(g, j, i)]);
      b = 0.5 * (w.muts0[i2(g, j, i - 1)] + w.muts0[i2(g, j, i)]);
    }
    w.muu[idx] = a;
    w.muus[idx] = b;
  } else if (idx < nu + nv) {
    size_t t = idx - nu;
    int i = (int)(t % (size_t)nx), j = (int)(t / (size_t)nx);
    double a, b;
    if (j == 0) {
      a = w.mut[i2(g, 0, i)];
      b = w.muts0[i2(g, 0, i)];
    } else if (j == ny) {
      a = w.mut[i2(g, ny - 1, i)];
      b = w.muts0[i2(g, ny - 1, i)];
    } else {
      a = 0.5 * (w.mut[i2(g, j - 1, i)] + w.mut[i2(g, j, i)]);
      b = 0.5 * (w.muts0[i2(g, j - 1, i)] + w.muts0[i2(g, j, i)]);
    }
    w.muv[t] = a;
    w.muvs[t] = b;
  }
}

__global__ void k_prep_u(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.nz * g.ny * (g.nx + 1);
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, g.nx + 1, g.ny, k, j, i);
  size_t f = i2u(g, j, i);
  double mref = g.c1h[k] * w.muus[f] + g.c2h[k];
  double mcur = g.c1h[k] * w.muu[f] + g.c2h[k];
  w.u_work[idx] = (mref * w.u0[idx] - mcur * s.u[idx]) / g.msfuy[f];
}

__global__ void k_prep_v(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.nz * (g.ny + 1) * g.nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, g.nx, g.ny + 1, k, j, i);
  size_t f = i2v(g, j, i);
  double mref = g.c1h[k] * w.muvs[f] + g.c2h[k];
  double mcur = g.c1h[k] * w.muv[f] + g.c2h[k];
  w.v_work[idx] = (mref * w.v0[idx] - mcur * s.v[idx]) / g.msfvx[f];
}

// w-level prep: w_work, ph_work, phb; the acoustic omega starts at ww_save.
__global__ void k_prep_w(DGrid g, DState s, DWork w) {
  size_t n = (size_t)(g.nz + 1) * g.ny * g.nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int k, j, i;
  decompose(idx, g.nx, g.ny, k, j, i);
  size_t c = i2(g, j, i);
  double mref = g.c1f[k] * w.muts0[c] + g.c2f[k];
  double mcur = g.c1f[k] * w.mut[c] + g.c2f[k];
  w.w_work[idx] = (mref * w.w0[idx] - mcur * s.w[idx]) / g.msfty[c];
  w.ph_work[idx] = w.ph0[idx] - s.ph_pert[idx];
  w.phb[idx] = s.ph_total[idx] - s.ph_pert[idx];
  w.ww[idx] = w.rom[idx];
}

// mass-level prep + calc_p_rho_wrf step 0 (core/calc_p_rho.py:48-139).
__global__ void k_prep_m(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.nz * g.ny * g.nx;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k, j, i;
  decompose(q, g.nx, g.ny, k, j, i);
  size_t c = i2(g, j, i), q1 = im(g, k + 1, j, i);
  double mref = g.c1h[k] * w.muts0[c] + g.c2h[k];
  double mcur = g.c1h[k] * w.mut[c] + g.c2h[k];
  double th_ref = w.th0[q] - kT0, th_cur = s.theta[q] - kT0;
  double th_work = mref * th_ref - mcur * th_cur;
  w.th_work[q] = th_work;
  double pb = s.p_total[q] - s.p_pert[q];
  w.pb[q] = pb;
  double alt_raw = w.alt_raw[q];
  double alt_ac = alt_raw > kMinAcousticAlt ? alt_raw : kMinAcousticAlt;
  w.alt[q] = alt_ac;
  double c2a = (kCpovcv * (pb + s.p_pert[q])) / alt_ac;
  w.c2a[q] = c2a;
  double phb_k = s.ph_total[q] - s.ph_pert[q];
  double phb_k1 = s.ph_total[q1] - s.ph_pert[q1];
  w.php[q] = 0.5 * (((phb_k + phb_k1) + s.ph_pert[q]) + s.ph_pert[q1]);
  // calc_p_rho step 0
  double mass_h = g.c1h[k] * w.muts0[c] + g.c2h[k];
  double safe_mass = safe_pos(mass_h, 1.0e-12);
  double mu_term = g.c1h[k] * w.mu_work0[c];
  double phw_k = w.ph0[q] - s.ph_pert[q];
  double phw_k1 = w.ph0[q1] - s.ph_pert[q1];
  double al = -(alt_ac * mu_term + g.rdnw[k] * (phw_k1 - phw_k)) / safe_mass;
  double ttr = kT0 + th_ref;
  double safe_t = safe_pos(ttr, 1.0e-6);
  double p = c2a * (((alt_ac * (th_work - mu_term * th_ref)) / (safe_mass * safe_t)) - al);
  w.al[q] = al;
  w.p[q] = p;
  w.pm1[q] = p;
  w.t2ave[q] = 0.0;
}

// ---------------------------------------------------------------------------
// Stage column setup (one thread per column):
//  rw_tend = pg_buoy_w_moist(qtot=0) + advect_w + curvature
//    (core/advance_w.py:209-290; operational_mode.py:1730-1822)
//  ph_tend = rhs_ph_wrf order 2 periodic, rigid lid (core/rhs_ph.py:175-398)
//  a/alpha/gamma = calc_coef_w_wrf_coefficients (dynamics/acoustic_wrf.py:783-852)
// ---------------------------------------------------------------------------
__global__ void k_col_setup(DGrid g, DNamelist nl, DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= nx) return;
  size_t c = i2(g, j, i);
  const double msfty = g.msfty[c];
  const double msft_inv = 1.0 / msfty;
  const double mut = w.mut[c], mub = w.mub[c];
  const double mu_p = mut - mub;
  // ---- rw_tend ----
  w.rw_tend[im(g, 0, j, i)] = 0.0 + w.wt[im(g, 0, j, i)];
  {
    double mxy = g.msftx[c] / msfty;
    for (int k = 1; k < nz; ++k) {
      double cq1 = 1.0 / (1.0 + 0.0), cq2 = 0.0 * cq1;
      double rw = (msft_inv * kG) *
                  (((cq1 * g.rdn[k]) * (s.p_pert[im(g, k, j, i)] - s.p_pert[im(g, k - 1, j, i)]) -
                    g.c1f[k] * mu_p) -
                   cq2 * (g.c1f[k] * mub + g.c2f[k]));
      rw = rw + w.wt[im(g, k, j, i)];
      double ru_m[2], u_m[2], rv_m[2], v_m[2];
      for (int sft = 0; sft < 2; ++sft) {
        int kk = k - sft;
        double mu_l = g.c1h[kk] * w.muu[i2u(g, j, i)] + g.c2h[kk];
        double mu_r = g.c1h[kk] * w.muu[i2u(g, j, i + 1)] + g.c2h[kk];
        double ru_l = (mu_l * s.u[iu(g, kk, j, i)]) / g.msfuy[i2u(g, j, i)];
        double ru_r = (mu_r * s.u[iu(g, kk, j, i + 1)]) / g.msfuy[i2u(g, j, i + 1)];
        ru_m[sft] = 0.5 * (ru_l + ru_r);
        u_m[sft] = 0.5 * (s.u[iu(g, kk, j, i)] + s.u[iu(g, kk, j, i + 1)]);
        double mv_s = g.c1h[kk] * w.muv[i2v(g, j, i)] + g.c2h[kk];
        double mv_n = g.c1h[kk] * w.muv[i2v(g, j + 1, i)] + g.c2h[kk];
        double rv_s = (mv_s * s.v[iv(g, kk, j, i)]) / g.msfvx[i2v(g, j, i)];
        double rv_n = (mv_n * s.v[iv(g, kk, j + 1, i)]) / g.msfvx[i2v(g, j + 1, i)];
        rv_m[sft] = 0.5 * (rv_s + rv_n);
        v_m[sft] = 0.5 * (s.v[iv(g, kk, j, i)] + s.v[iv(g, kk, j + 1, i)]);
      }
      double ru_f = g.fnm[k] * ru_m[0] + g.fnp[k] * ru_m[1];
      double u_f = g.fnm[k] * u_m[0] + g.fnp[k] * u_m[1];
      double rv_f = g.fnm[k] * rv_m[0] + g.fnp[k] * rv_m[1];
      double v_f = g.fnm[k] * v_m[0] + g.fnp[k] * v_m[1];
      double curv = kReradius * (ru_f * u_f + (mxy * rv_f) * v_f);
      w.rw_tend[im(g, k, j, i)] = rw + (0.0 + curv);
    }
    double cq1 = 1.0 / (1.0 + 0.0), cq2 = 0.0 * cq1;
    double rwt = (msft_inv * kG) *
                 (((((cq1 * 2.0) * g.rdnw[nz - 1]) * (-s.p_pert[im(g, nz - 1, j, i)])) - g.c1f[nz] * mu_p) -
                  cq2 * (g.c1f[nz] * mub + g.c2f[nz]));
    w.rw_tend[im(g, nz, j, i)] = rwt + w.wt[im(g, nz, j, i)];
  }
  // ---- ph_tend (rhs_ph) ----
  {
    int jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny), ip = wrapi(i + 1, nx), imn = wrapi(i - 1, nx);
    auto PHT = [&](int k, int jj, int ii) { return s.ph_pert[im(g, k, jj, ii)] + w.phb[im(g, k, jj, ii)]; };
    w.ph_tend[im(g, 0, j, i)] = 0.0;
    for (int f = 1; f < nz; ++f) {
      double wd_f = ((0.5 * (w.rom[im(g, f + 1, j, i)] + w.rom[im(g, f, j, i)])) * g.rdnw[f]) *
                    (PHT(f + 1, j, i) - PHT(f, j, i));
      double wd_fm = ((0.5 * (w.rom[im(g, f, j, i)] + w.rom[im(g, f - 1, j, i)])) * g.rdnw[f - 1]) *
                     (PHT(f, j, i) - PHT(f - 1, j, i));
      double term3 = g.fnm[f] * wd_f + g.fnp[f] * wd_fm;
      double t = -term3;
      double mass_f = g.c1f[f] * mut + g.c2f[f];
      double gw = ((mass_f * kG) * s.w[im(g, f, j, i)]) * msft_inv;
      t = t + gw;
      double pti = PHT(f, j, i);
      double dn_ = PHT(f, jp, i) - pti, ds_ = pti - PHT(f, jm, i);
      double vn = s.v[iv(g, f, j + 1, i)] + s.v[iv(g, f - 1, j + 1, i)];
      double vs = s.v[iv(g, f, j, i)] + s.v[iv(g, f - 1, j, i)];
      double mvn = g.c1f[f] * w.muv[i2v(g, j + 1, i)] + g.c2f[f];
      double mvs = g.c1f[f] * w.muv[i2v(g, j, i)] + g.c2f[f];
      double fy = (mvn * vn) * dn_ + (mvs * vs) * ds_;
      t = t + (-((0.25 * g.rdy) * msft_inv)) * fy;
      double de_ = PHT(f, j, ip) - pti, dw_ = pti - PHT(f, j, imn);
      double ue = s.u[iu(g, f, j, i + 1)] + s.u[iu(g, f - 1, j, i + 1)];
      double uw = s.u[iu(g, f, j, i)] + s.u[iu(g, f - 1, j, i)];
      double mue = g.c1f[f] * w.muu[i2u(g, j, i + 1)] + g.c2f[f];
      double muw = g.c1f[f] * w.muu[i2u(g, j, i)] + g.c2f[f];
      double fx = (mue * ue) * de_ + (muw * uw) * dw_;
      t = t + (-((0.25 * g.rdx) * msft_inv)) * fx;
      w.ph_tend[im(g, f, j, i)] = t;
    }
    w.ph_tend[im(g, nz, j, i)] = 0.0;
  }
  // ---- calc_coef_w ----
  {
    double cof = (0.5 * dts * kG * (1.0 + nl.epssm));
    cof = cof * cof;
    double lid_flag = nl.top_lid ? 0.0 : 1.0;
    auto MH = [&](int k) { return g.c1h[k] * mut + g.c2h[k]; };
    auto MF = [&](int k) { return g.c1f[k] * mut + g.c2f[k]; };
    auto C2A = [&](int k) { return w.c2a[im(g, k, j, i)]; };
    double top_denom_a = MH(nz - 1) * MF(nz - 1);
    double top_denom_b = MH(nz - 1) * MF(nz);
    double r2 = g.rdnw[nz - 1] * g.rdnw[nz - 1];
    double a_top = ((((-2.0 * cof) * r2) * C2A(nz - 1)) * lid_flag) / top_denom_a;
    w.ca[im(g, 0, j, i)] = 0.0;
    w.calpha[im(g, 0, j, i)] = 1.0;
    w.cgamma[im(g, 0, j, i)] = 0.0;
    double gamma_prev = 0.0;
    for (int k = 1; k < nz; ++k) {
      double ak;
      if (k == 1) {
        ak = 0.0;
      } else {
        int kb = k - 1;
        double denom = MH(kb) * MF(kb);
        ak = (((((-cqw_at(k, nz)) * cof) * g.rdn[k]) * g.rdnw[k - 1]) * C2A(k - 1)) / denom;
      }
      double du = MH(k) * MF(k);
      double dl = MH(k - 1) * MF(k);
      double dc = MH(k) * MF(k + 1);
      double b = 1.0 + ((cqw_at(k, nz) * cof) * g.rdn[k]) *
                           ((g.rdnw[k] * C2A(k)) / du + (g.rdnw[k - 1] * C2A(k - 1)) / dl);
      double cc = (((((-cqw_at(k, nz)) * cof) * g.rdn[k]) * g.rdnw[k]) * C2A(k)) / dc;
      double al = 1.0 / (b - ak * gamma_prev);
      double gm = cc * al;
      w.ca[im(g, k, j, i)] = ak;
      w.calpha[im(g, k, j, i)] = al;
      w.cgamma[im(g, k, j, i)] = gm;
      gamma_prev = gm;
    }
    double b_top = 1.0 + (((2.0 * cof) * r2) * C2A(nz - 1)) / top_denom_b;
    w.ca[im(g, nz, j, i)] = a_top;
    w.calpha[im(g, nz, j, i)] = 1.0 / (b_top - a_top * gamma_prev);
    w.cgamma[im(g, nz, j, i)] = 0.0;
  }
}

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_stage_setup(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              int rk_step, double dts, hipStream_t st, Prof* prof) {
  const unsigned bs = 256;
  const size_t M = (size_t)g.nz * g.ny * g.nx;
  const size_t Wn = (size_t)(g.nz + 1) * g.ny * g.nx;
  const size_t U = (size_t)g.nz * g.ny * (g.nx + 1);
  const size_t V = (size_t)g.nz * (g.ny + 1) * g.nx;
  const size_t C = (size_t)g.ny * g.nx;
  const size_t F = (size_t)g.ny * (g.nx + 1) + (size_t)(g.ny + 1) * g.nx;
  dim3 colgrid((g.nx + 63) / 64, g.ny);
  WD_TIMED(prof, KID_RUV, st, (k_ruv<<<blocks_for(M, bs), bs, 0, st>>>(g, s.u, s.v, s.mu_total, w.ru, w.rv)));
  WD_TIMED(prof, KID_ROM, st, (k_rom<<<colgrid, 64, 0, st>>>(g, w.ru, w.rv, w.rom)));
  WD_TIMED(prof, KID_DIAG, st, (k_diag<<<blocks_for(M, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_TEND_U, st, (k_tend_u<<<blocks_for(U, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_TEND_V, st, (k_tend_v<<<blocks_for(V, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_TEND_W, st, (k_tend_w<<<blocks_for(Wn, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_TEND_T, st, (k_tend_t<<<blocks_for(M, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_PREP_2D, st, (k_prep_2d<<<blocks_for(C, bs), bs, 0, st>>>(g, s, w, rk_step)));
  WD_TIMED(prof, KID_PREP_FACES, st, (k_prep_faces<<<blocks_for(F, bs), bs, 0, st>>>(g, w)));
  WD_TIMED(prof, KID_PREP_U, st, (k_prep_u<<<blocks_for(U, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_PREP_V, st, (k_prep_v<<<blocks_for(V, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_PREP_W, st, (k_prep_w<<<blocks_for(Wn, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_PREP_M, st, (k_prep_m<<<blocks_for(M, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_COL_SETUP, st, (k_col_setup<<<colgrid, 64, 0, st>>>(g, nl, s, w, dts)));
  return hipGetLastError();
}

}  // namespace wd
