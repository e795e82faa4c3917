// Acoustic small step (advance_uv, then the fused column kernel
// advance_mu_t + advance_w + calc_p_rho) and the stage finish
// (small_step_finish + grid%p refresh).
//
// Each output value is computed with the reference's expression order; build
// with -ffp-contract=off.
#include "wrfdyn_kernels.h"

namespace wd {

namespace {

__device__ inline void decompose(size_t idx, int nxx, int nyy, int& k, int& j, int& i) {
  i = (int)(idx % (size_t)nxx);
  size_t t = idx / (size_t)nxx;
  j = (int)(t % (size_t)nyy);
  k = (int)(t / (size_t)nyy);
}

// dpn face value for the acoustic PGF (core/acoustic.py:582-623): pair sums of
// the work pressure; the rigid-lid top face uses cf1/cf2/cf3 on the top three levels.
__device__ inline double dpn_ac(const DGrid& g, const double* __restrict__ p, int kk, size_t ca, size_t cb,
                                size_t plane) {
#define PS(k) (p[(size_t)(k) * plane + ca] + p[(size_t)(k) * plane + cb])
  if (kk == 0) return 0.5 * ((g.cf1 * PS(0) + g.cf2 * PS(1)) + g.cf3 * PS(2));
  if (kk == g.nz) return 0.5 * ((g.cf1 * PS(g.nz - 1) + g.cf2 * PS(g.nz - 2)) + g.cf3 * PS(g.nz - 3));
  return 0.5 * (g.fnm[kk] * PS(kk) + g.fnp[kk] * PS(kk - 1));
#undef PS
}

// ---------------------------------------------------------------------------
// advance_uv_wrf, core/acoustic.py:626-786 (emdiv = 0.01), u faces.
// ---------------------------------------------------------------------------
__global__ void k_uv_u(DGrid g, DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)nz * ny * (nx + 1);
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k, j, i;
  decompose(q, nx + 1, ny, k, j, i);
  int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
  size_t cl = i2(g, j, il), cr = i2(g, j, ir);
  size_t plane = (size_t)ny * nx;
  double msf = g.msfux[i2u(g, j, i)] / g.msfuy[i2u(g, j, i)];
  double muw_l = w.muts[cl] - w.mut[cl];
  double muw_r = w.muts[cr] - w.mut[cr];
  double dpn_k = dpn_ac(g, w.p, k, cl, cr, plane);
  double dpn_k1 = dpn_ac(g, w.p, k + 1, cl, cr, plane);
  double mudf_xy = ((-kEmdiv * g.dx) * (s.mudf[cr] - s.mudf[cl])) / g.msfuy[i2u(g, j, i)];
  double u = w.u_work[q] + dts * w.ut[q];
  double mass_x = g.c1h[k] * w.muu[i2u(g, j, i)] + g.c2h[k];
  size_t ql = im(g, k, j, il), qr = im(g, k, j, ir), ql1 = im(g, k + 1, j, il), qr1 = im(g, k + 1, j, ir);
  double ph_term = (w.ph_work[qr1] - w.ph_work[ql1]) + (w.ph_work[qr] - w.ph_work[ql]);
  double p_term = (w.alt[ql] + w.alt[qr]) * (w.p[qr] - w.p[ql]);
  double pb_term = (w.al[ql] + w.al[qr]) * (w.pb[qr] - w.pb[ql]);
  double dpx = (((msf * 0.5) * g.rdx) * mass_x) * ((ph_term + p_term) + pb_term);
  double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (muw_l + muw_r));
  dpx = dpx + ((msf * g.rdx) * (w.php[qr] - w.php[ql])) * bracket;
  u = u - (dts * 1.0) * dpx;
  u = u + g.c1h[k] * mudf_xy;
  w.u_work[q] = u;
}

// advance_uv_wrf v faces.
__global__ void k_uv_v(DGrid g, DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  size_t n = (size_t)nz * (ny + 1) * nx;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k, j, i;
  decompose(q, nx, ny + 1, k, j, i);
  int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
  size_t cs = i2(g, js, i), cn = i2(g, jn, i);
  size_t plane = (size_t)ny * nx;
  double msf = g.msfvy[i2v(g, j, i)] / g.msfvx[i2v(g, j, i)];
  double muw_s = w.muts[cs] - w.mut[cs];
  double muw_n = w.muts[cn] - w.mut[cn];
  double dpn_k = dpn_ac(g, w.p, k, cs, cn, plane);
  double dpn_k1 = dpn_ac(g, w.p, k + 1, cs, cn, plane);
  double msfvx_inv = 1.0 / g.msfvx[i2v(g, j, i)];
  double mudf_xy = ((-kEmdiv * g.dy) * (s.mudf[cn] - s.mudf[cs])) * msfvx_inv;
  double v = w.v_work[q] + dts * w.vt[q];
  double mass_y = g.c1h[k] * w.muv[i2v(g, j, i)] + g.c2h[k];
  size_t qs = im(g, k, js, i), qn = im(g, k, jn, i), qs1 = im(g, k + 1, js, i), qn1 = im(g, k + 1, jn, i);
  double ph_term = (w.ph_work[qn1] - w.ph_work[qs1]) + (w.ph_work[qn] - w.ph_work[qs]);
  double p_term = (w.alt[qs] + w.alt[qn]) * (w.p[qn] - w.p[qs]);
  double pb_term = (w.al[qs] + w.al[qn]) * (w.pb[qn] - w.pb[qs]);
  double dpy = (((msf * 0.5) * g.rdy) * mass_y) * ((ph_term + p_term) + pb_term);
  double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (muw_s + muw_n));
  dpy = dpy + ((msf * g.rdy) * (w.php[qn] - w.php[qs])) * bracket;
  v = v - (dts * 1.0) * dpy;
  v = v + g.c1h[k] * mudf_xy;
  w.v_work[q] = v;
}

// ---------------------------------------------------------------------------
// Fused column kernel, one thread per (j, i) column:
//   advance_mu_t periodic   dynamics/mu_t_advance.py:139-292
//   advance_w_wrf           core/advance_w.py:293-627 (w_damp :92-132,
//                           Thomas sweeps :558-592, damp_opt=3 :594-617)
//   calc_p_rho_step         core/calc_p_rho.py:142-183 (smdiv = 0.1)
// ---------------------------------------------------------------------------
__global__ void k_column(DGrid g, DNamelist nl, DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= nx) return;
  const size_t c = i2(g, j, i);
  const double msftx = g.msftx[c], msfty = g.msfty[c];
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double mut = w.mut[c];

  // ---- advance_mu_t: dvdxi and dmdt ----
  double dmdt = 0.0;
  for (int k = 0; k < nz; ++k) {
    double c1 = g.c1h[k], c2 = g.c2h[k];
    double v_n = w.v_work[iv(g, k, j + 1, i)] +
                 ((c1 * w.muv[i2v(g, j + 1, i)] + c2) * s.v[iv(g, k, j + 1, i)]) * (1.0 / g.msfvx[i2v(g, j + 1, i)]);
    double v_s = w.v_work[iv(g, k, j, i)] +
                 ((c1 * w.muv[i2v(g, j, i)] + c2) * s.v[iv(g, k, j, i)]) * (1.0 / g.msfvx[i2v(g, j, i)]);
    double u_e = w.u_work[iu(g, k, j, i + 1)] +
                 ((c1 * w.muu[i2u(g, j, i + 1)] + c2) * s.u[iu(g, k, j, i + 1)]) / g.msfuy[i2u(g, j, i + 1)];
    double u_w = w.u_work[iu(g, k, j, i)] +
                 ((c1 * w.muu[i2u(g, j, i)] + c2) * s.u[iu(g, k, j, i)]) / g.msfuy[i2u(g, j, i)];
    double d = (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
    w.dvdxi[im(g, k, j, i)] = d;
    dmdt = (k == 0) ? g.dnw[0] * d : dmdt + g.dnw[k] * d;
  }
  const double mu_work_old = w.muts[c] - mut;
  const double mu_save = w.mu[c] - mu_work_old;
  const double mu_tendency = dmdt + 0.0;
  const double mu_work_new = mu_work_old + dts * mu_tendency;
  w.mu[c] = mu_save + mu_work_new;
  s.mudf[c] = mu_tendency;
  const double muts = mut + mu_work_new;
  w.muts[c] = muts;
  const double muave = 0.5 * (eps_p * mu_work_new + eps_m * mu_work_old);
  w.muave[c] = muave;

  // ---- ww recurrence (ww_1 = ww_save = rom) and the coupled theta update ----
  {
    int ip = wrapi(i + 1, nx), imn = wrapi(i - 1, nx), jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny);
    auto T1 = [&](int k, int jj, int ii) { return s.theta[im(g, k, jj, ii)] - kT0; };
    double wprev = w.ww[im(g, 0, j, i)];
    w.ww[im(g, 0, j, i)] = wprev - w.rom[im(g, 0, j, i)];
    double wd_lo = 0.0;
    for (int k = 0; k < nz; ++k) {
      double wd_hi;
      if (k + 1 < nz) {
        int kk = k + 1;
        double incr = (g.dnw[kk - 1] * ((g.c1h[k] * dmdt + w.dvdxi[im(g, kk - 1, j, i)]) + g.c1h[k] * 0.0)) / msfty;
        double wk = wprev - incr;
        double wwk = wk - w.rom[im(g, kk, j, i)];
        w.ww[im(g, kk, j, i)] = wwk;
        wprev = wk;
        double face = g.fnm[kk] * T1(kk, j, i) + g.fnp[kk] * T1(kk - 1, j, i);
        wd_hi = wwk * face;
      } else {
        wd_hi = 0.0;
      }
      double th = T1(k, j, i);
      double v_flux = w.v_work[iv(g, k, j + 1, i)] * (T1(k, jp, i) + th) - w.v_work[iv(g, k, j, i)] * (th + T1(k, jm, i));
      double u_flux = w.u_work[iu(g, k, j, i + 1)] * (T1(k, j, ip) + th) - w.u_work[iu(g, k, j, i)] * (th + T1(k, j, imn));
      double tendency = msftx * ((0.5 * g.rdy) * v_flux + (0.5 * g.rdx) * u_flux) + g.rdnw[k] * (wd_hi - wd_lo);
      size_t q = im(g, k, j, i);
      double theta_i = w.th_work[q] + ((msfty * dts) * w.tht[q]);
      w.th_work[q] = theta_i - ((dts * msfty) * tendency);
      wd_lo = wd_hi;
    }
    // ww[nz] keeps its value (not offset by ww_1)
  }

  // ---- advance_w ----
  const double msft_inv = 1.0 / msfty;
  // t_2ave (stored in place once the old value is consumed)
  // rhs of the phi equation into the column scratch
  double* rhs = w.rhs;
  auto RH = [&](int f) -> double& { return rhs[im(g, f, j, i)]; };
  auto PH1T = [&](int k) { return w.ph0[im(g, k, j, i)] + w.phb[im(g, k, j, i)]; };
  {
    const double c05geps = (0.5 * kG) * eps_m;
    RH(0) = 0.0;
    for (int f = 1; f <= nz; ++f) {
      double r = dts * (w.ph_tend[im(g, f, j, i)] + c05geps * w.w_work[im(g, f, j, i)]);
      if (f < nz) {
        double wd_f1 = ((0.5 * (w.ww[im(g, f + 1, j, i)] + w.ww[im(g, f, j, i)])) * g.rdnw[f]) * (PH1T(f + 1) - PH1T(f));
        double wd_f = ((0.5 * (w.ww[im(g, f, j, i)] + w.ww[im(g, f - 1, j, i)])) * g.rdnw[f - 1]) * (PH1T(f) - PH1T(f - 1));
        double adv = dts * (g.fnm[f] * wd_f1 + g.fnp[f] * wd_f);
        r = r + (-adv);
      }
      double mass_f = g.c1f[f] * mut + g.c2f[f];
      r = w.ph_work[im(g, f, j, i)] + (msfty * r) / mass_f;
      if (f == nz && nl.top_lid) r = 0.0;
      RH(f) = r;
    }
  }
  // surface w from terrain (edge-padded ht differences)
  double w_surface;
  {
    int jn = clampi(j + 1, 0, ny - 1), js = clampi(j - 1, 0, ny - 1);
    int ie = clampi(i + 1, 0, nx - 1), iw = clampi(i - 1, 0, nx - 1);
    double ht = w.phb[im(g, 0, j, i)] / kG;
    double ht_n = w.phb[im(g, 0, jn, i)] / kG, ht_s = w.phb[im(g, 0, js, i)] / kG;
    double ht_e = w.phb[im(g, 0, j, ie)] / kG, ht_w = w.phb[im(g, 0, j, iw)] / kG;
    double dy_n = ht_n - ht, dy_s = ht - ht_s, dx_e = ht_e - ht, dx_w = ht - ht_w;
    double v_cf_n = (g.cf1 * s.v[iv(g, 0, j + 1, i)] + g.cf2 * s.v[iv(g, 1, j + 1, i)]) + g.cf3 * s.v[iv(g, 2, j + 1, i)];
    double v_cf_s = (g.cf1 * s.v[iv(g, 0, j, i)] + g.cf2 * s.v[iv(g, 1, j, i)]) + g.cf3 * s.v[iv(g, 2, j, i)];
    double u_cf_e = (g.cf1 * s.u[iu(g, 0, j, i + 1)] + g.cf2 * s.u[iu(g, 1, j, i + 1)]) + g.cf3 * s.u[iu(g, 2, j, i + 1)];
    double u_cf_w = (g.cf1 * s.u[iu(g, 0, j, i)] + g.cf2 * s.u[iu(g, 1, j, i)]) + g.cf3 * s.u[iu(g, 2, j, i)];
    w_surface = ((msfty * 0.5) * g.rdy) * (dy_n * v_cf_n + dy_s * v_cf_s) +
                ((msftx * 0.5) * g.rdx) * (dx_e * u_cf_e + dx_w * u_cf_w);
  }
  // t_2ave for this substep (levels k) -- computed on the fly where needed
  auto T2N = [&](int k) {
    size_t q = im(g, k, j, i);
    double mass_h = g.c1h[k] * muts + g.c2h[k];
    double half = 0.5 * (eps_p * w.th_work[q] + eps_m * w.t2ave[q]);
    double ttr = kT0 + (s.theta[q] - kT0);
    return (half + (g.c1h[k] * muave) * kT0) / (mass_h * ttr);
  };
  // explicit w + forward Thomas sweep; w_work holds the forward sweep
  {
    const double c05dtg = (0.5 * dts) * kG;
    double wfwd_prev = w_surface;
    double t2_lo = T2N(0);
    for (int f = 1; f <= nz; ++f) {
      size_t qf = im(g, f, j, i);
      double wold = w.w_work[qf];
      double rw = w.rw_tend[qf];
      if (nl.w_damping == 1 && f < nz) {
        double mass_f = g.c1f[f] * mut + g.c2f[f];
        double safe_mf = safe_pos(mass_f, 1.0e-12);
        double vert_cfl = fabs(((w.ww[qf] / safe_mf) * g.rdnw[f]) * dts);
        double damp = (((-sgn(wold)) * kWAlpha) * (vert_cfl - kWBeta)) * mass_f;
        rw = rw + (vert_cfl > kWBeta ? damp : 0.0);
      }
      double wn = wold + dts * rw;
      double t2_hi = 0.0;
      if (f < nz) {
        t2_hi = T2N(f);
        double mh_u = safe_pos(g.c1h[f] * mut + g.c2h[f], 1.0e-12);
        double mh_l = safe_pos(g.c1h[f - 1] * mut + g.c2h[f - 1], 1.0e-12);
        double coef_u = (w.c2a[im(g, f, j, i)] * g.rdnw[f]) / mh_u;
        double coef_l = (w.c2a[im(g, f - 1, j, i)] * g.rdnw[f - 1]) / mh_l;
        double phk1 = w.ph_work[im(g, f + 1, j, i)], phk = w.ph_work[qf], phkm = w.ph_work[im(g, f - 1, j, i)];
        double tA_u = coef_u * (eps_p * (RH(f + 1) - RH(f)) + eps_m * (phk1 - phk));
        double tA_l = coef_l * (eps_p * (RH(f) - RH(f - 1)) + eps_m * (phk - phkm));
        double termA = ((msft_inv * cqw_at(f, nz)) * (c05dtg * g.rdn[f])) * (tA_u - tA_l);
        double bu = (w.c2a[im(g, f, j, i)] * w.alt[im(g, f, j, i)]) * t2_hi;
        double bl = (w.c2a[im(g, f - 1, j, i)] * w.alt[im(g, f - 1, j, i)]) * t2_lo;
        double termB = ((dts * kG) * msft_inv) * (g.rdn[f] * (bu - bl) - (g.c1f[f] * muave));
        wn = wn + (termA + termB);
      } else {
        wn = 0.0;  // rigid lid
      }
      double wf = (wn - w.ca[qf] * wfwd_prev) * w.calpha[qf];
      w.w_work[qf] = wf;
      wfwd_prev = wf;
      if (f < nz) {
        w.t2ave[im(g, f - 1, j, i)] = t2_lo;
        t2_lo = t2_hi;
      } else {
        w.t2ave[im(g, nz - 1, j, i)] = t2_lo;
      }
    }
    w.w_work[im(g, 0, j, i)] = w_surface;
  }
  // back substitution, damping, geopotential finish (top-down)
  {
    const bool damp = (nl.damp_opt == 3 && nl.dampcoef > 0.0);
    const double dampmag = dts * nl.dampcoef, hdepth = nl.zdamp;
    const double htop = PH1T(nz) / kG;
    const double hbot = htop - hdepth;
    double wnext = w.w_work[im(g, nz, j, i)];
    for (int f = nz; f >= 1; --f) {
      size_t qf = im(g, f, j, i);
      double wv;
      if (f == nz) {
        wv = wnext;
      } else {
        wv = w.w_work[qf] - w.cgamma[qf] * wnext;
        wnext = wv;
      }
      double wd = wv;
      if (damp) {
        double hk = PH1T(f) / kG;
        double sn = wd_sin(((0.5 * kPi) * (hk - hbot)) / hdepth);
        double ramp = sn * sn;
        double dampwt = hk >= hbot ? dampmag * ramp : 0.0;
        double mass_f = g.c1f[f] * mut + g.c2f[f];
        wd = (wv - (dampwt * mass_f) * s.w[qf]) / (1.0 + dampwt);
      }
      double mfs = g.c1f[f] * muts + g.c2f[f];
      w.ph_work[qf] = RH(f) + (((((msfty * 0.5) * dts) * kG) * eps_p) * wd) / mfs;
      w.w_work[qf] = wd;
    }
  }
  // ---- calc_p_rho_step ----
  {
    const double mu_work = muts - mut;
    for (int k = 0; k < nz; ++k) {
      size_t q = im(g, k, j, i);
      double mass_h = g.c1h[k] * muts + g.c2h[k];
      double safe_mass = safe_pos(mass_h, 1.0e-12);
      double mu_term = g.c1h[k] * mu_work;
      double al = -(w.alt[q] * mu_term + g.rdnw[k] * (w.ph_work[im(g, k + 1, j, i)] - w.ph_work[q])) / safe_mass;
      double theta_1 = s.theta[q] - kT0;
      double ttr = kT0 + theta_1;
      double safe_t = safe_pos(ttr, 1.0e-6);
      double p = w.c2a[q] * (((w.alt[q] * (w.th_work[q] - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
      double ptmp = p;
      p = p + kSmdiv * (p - w.pm1[q]);
      w.al[q] = al;
      w.p[q] = p;
      w.pm1[q] = ptmp;
    }
  }
}

// ---------------------------------------------------------------------------
// small_step_finish_wrf (core/small_step_finish.py:22-101)
// ---------------------------------------------------------------------------
__device__ inline double safe_den(double ms) { return fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12); }

__global__ void k_fin_w(DGrid g, DState s, DWork w) {
  size_t n = (size_t)(g.nz + 1) * g.ny * g.nx;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k, j, i;
  decompose(q, g.nx, g.ny, k, j, i);
  size_t c = i2(g, j, i);
  double ms = g.c1f[k] * w.muts[c] + g.c2f[k];
  double mc = g.c1f[k] * w.mut[c] + g.c2f[k];
  s.w[q] = (g.msfty[c] * w.w_work[q] + s.w[q] * mc) / safe_den(ms);
  double ph_pert = w.ph_work[q] + s.ph_pert[q];
  double ph_base = s.ph_total[q] - s.ph_pert[q];
  s.ph_total[q] = ph_base + ph_pert;
  s.ph_pert[q] = ph_pert;
}

__global__ void k_fin_uv(DGrid g, DState s, DWork w) {
  size_t nu = (size_t)g.nz * g.ny * (g.nx + 1);
  size_t nv = (size_t)g.nz * (g.ny + 1) * g.nx;
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k, j, i;
  if (idx < nu) {
    decompose(idx, g.nx + 1, g.ny, k, j, i);
    size_t f = i2u(g, j, i);
    double ms = g.c1h[k] * w.muus[f] + g.c2h[k];
    double mc = g.c1h[k] * w.muu[f] + g.c2h[k];
    s.u[idx] = (g.msfuy[f] * w.u_work[idx] + s.u[idx] * mc) / safe_den(ms);
  } else if (idx < nu + nv) {
    size_t q = idx - nu;
    decompose(q, g.nx, g.ny + 1, k, j, i);
    size_t f = i2v(g, j, i);
    double ms = g.c1h[k] * w.muvs[f] + g.c2h[k];
    double mc = g.c1h[k] * w.muv[f] + g.c2h[k];
    s.v[q] = (g.msfvx[f] * w.v_work[q] + s.v[q] * mc) / safe_den(ms);
  }
}

// theta, p (finish then _refresh_grid_p_from_finished, operational_mode.py:2169-2231
// with diagnose_pressure_al_alt's base branch, acoustic_wrf.py:360-438).
// Requires k_fin_w (new ph_total/ph_pert) to have run.
__global__ void k_fin_m(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.nz * g.ny * g.nx;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k, j, i;
  decompose(q, g.nx, g.ny, k, j, i);
  size_t c = i2(g, j, i), q1 = im(g, k + 1, j, i);
  double ms = g.c1h[k] * w.muts[c] + g.c2h[k];
  double mc = g.c1h[k] * w.mut[c] + g.c2h[k];
  double t_save = s.theta[q] - kT0;
  double theta_pert = (w.th_work[q] + t_save * mc) / safe_den(ms);
  double theta = theta_pert + kT0;
  s.theta[q] = theta;
  double p_base = s.p_total[q] - s.p_pert[q];
  double p_total = p_base + w.p[q];
  double p_pert = w.p[q];
  // refresh
  double mub = w.mub[c];
  double mu_pert_new = w.mu[c];
  double muts = mub + mu_pert_new;
  double phb_k = s.ph_total[q] - s.ph_pert[q];
  double phb_k1 = s.ph_total[q1] - s.ph_pert[q1];
  double mass_h_base = g.c1h[k] * mub + g.c2h[k];
  double denom_alb = g.dnw[k] * mass_h_base;
  double safe_da = safe_pos(denom_alb, 1.0e-12);
  double alb = -(phb_k1 - phb_k) / safe_da;
  double mass_w = g.c1h[k] * muts + g.c2h[k];
  double safe_mass = safe_pos(mass_w, 1.0e-12);
  double mu_term = (alb * g.c1h[k]) * mu_pert_new;
  double geo_term = g.rdnw[k] * (s.ph_pert[q1] - s.ph_pert[q]);
  double al = -(mu_term + geo_term) / safe_mass;
  double alt = al + alb;
  double ptot = p_from_theta_alt(theta, alt);
  double p_pert_new = ptot - w.pb[q];
  double p_base2 = p_total - p_pert;
  s.p_total[q] = p_base2 + p_pert_new;
  s.p_pert[q] = p_pert_new;
}

__global__ void k_fin_mu(DGrid g, DState s, DWork w) {
  size_t n = (size_t)g.ny * g.nx;
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double mu_base = s.mu_total[c] - s.mu_pert[c];
  s.mu_pert[c] = w.mu[c];
  s.mu_total[c] = mu_base + w.mu[c];
}

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_acoustic_substep(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                   double dts, hipStream_t st, Prof* prof) {
  const unsigned bs = 256;
  const size_t U = (size_t)g.nz * g.ny * (g.nx + 1);
  const size_t V = (size_t)g.nz * (g.ny + 1) * g.nx;
  WD_TIMED(prof, KID_UV_U, st, (k_uv_u<<<blocks_for(U, bs), bs, 0, st>>>(g, s, w, dts)));
  WD_TIMED(prof, KID_UV_V, st, (k_uv_v<<<blocks_for(V, bs), bs, 0, st>>>(g, s, w, dts)));
  dim3 colgrid((g.nx + 63) / 64, g.ny);
  WD_TIMED(prof, KID_COLUMN, st, (k_column<<<colgrid, 64, 0, st>>>(g, nl, s, w, dts)));
  return hipGetLastError();
}

hipError_t launch_stage_finish(const DGrid& g, const DState& s, const DWork& w, hipStream_t st, Prof* prof) {
  const unsigned bs = 256;
  const size_t M = (size_t)g.nz * g.ny * g.nx;
  const size_t Wn = (size_t)(g.nz + 1) * g.ny * g.nx;
  const size_t UV = (size_t)g.nz * g.ny * (g.nx + 1) + (size_t)g.nz * (g.ny + 1) * g.nx;
  const size_t C = (size_t)g.ny * g.nx;
  WD_TIMED(prof, KID_FIN_W, st, (k_fin_w<<<blocks_for(Wn, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_FIN_UV, st, (k_fin_uv<<<blocks_for(UV, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_FIN_M, st, (k_fin_m<<<blocks_for(M, bs), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_FIN_MU, st, (k_fin_mu<<<blocks_for(C, bs), bs, 0, st>>>(g, s, w)));
  return hipGetLastError();
}

const char* kernel_name(int id) {
  static const char* names[KID_COUNT] = {
      "origin_copy", "k_ruv", "k_rom", "k_diag", "k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t",
      "k_prep_2d", "k_prep_faces", "k_prep_u", "k_prep_v", "k_prep_w", "k_prep_m", "k_col_setup",
      "k_uv_u", "k_uv_v", "k_column", "k_fin_w", "k_fin_uv", "k_fin_m", "k_fin_mu"};
  return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

}  // namespace wd
