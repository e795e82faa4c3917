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
// advance_uv_wrf, core/acoustic.py:626-786 (emdiv = 0.01).
// One thread per (j, i) face column of the (ny+1) x (nx+1) face grid marches
// up k and advances u at face (j, i) when j < ny and v at face (j, i) when
// i < nx; the mass column (j, i) is shared by both stencils.  The dpn pair
// sums of the work pressure are carried up the column (interior dpn(k) uses
// levels k and k-1; the bottom and rigid-lid top faces use three levels).
// ---------------------------------------------------------------------------
template <int BX, int BY>
__global__ void __launch_bounds__(BX * BY) k_uv(DGrid g, DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  if (i > nx || j > ny) return;
  const bool do_u = j < ny;
  const bool do_v = i < nx;
  const size_t plane = (size_t)ny * nx;
  // mass columns: centre (clamped), west and south neighbours (edge padded)
  const int ic = clampi(i, 0, nx - 1), iw = clampi(i - 1, 0, nx - 1);
  const int jc = clampi(j, 0, ny - 1), js = clampi(j - 1, 0, ny - 1);
  const size_t cc = i2(g, jc, ic);
  const size_t cw = i2(g, jc, iw);  // u pair: (west, centre) on row jc == j
  const size_t cs = i2(g, js, ic);  // v pair: (south, centre) on column ic == i
  double msf_u = 0.0, mudf_u = 0.0, muwsum_u = 0.0, muu = 0.0;
  double msf_v = 0.0, mudf_v = 0.0, muwsum_v = 0.0, muv = 0.0;
  if (do_u) {
    msf_u = g.msfux[i2u(g, j, i)] / g.msfuy[i2u(g, j, i)];
    double muw_l = w.muts[cw] - w.mut[cw];
    double muw_r = w.muts[cc] - w.mut[cc];
    muwsum_u = muw_l + muw_r;
    mudf_u = ((-kEmdiv * g.dx) * (s.mudf[cc] - s.mudf[cw])) / g.msfuy[i2u(g, j, i)];
    muu = w.muu[i2u(g, j, i)];
  }
  if (do_v) {
    msf_v = g.msfvy[i2v(g, j, i)] / g.msfvx[i2v(g, j, i)];
    double muw_s = w.muts[cs] - w.mut[cs];
    double muw_n = w.muts[cc] - w.mut[cc];
    muwsum_v = muw_s + muw_n;
    double msfvx_inv = 1.0 / g.msfvx[i2v(g, j, i)];
    mudf_v = ((-kEmdiv * g.dy) * (s.mudf[cc] - s.mudf[cs])) * msfvx_inv;
    muv = w.muv[i2v(g, j, i)];
  }
  const double* P = w.p;
  // pair sums of p at levels k (carried); psu2/psu1/psu0 = levels k-2, k-1, k
  double pc0 = P[cc], pc1 = P[plane + cc], pc2 = P[2 * plane + cc];
  double pw0 = P[cw], pw1 = P[plane + cw], pw2 = P[2 * plane + cw];
  double ps0v = P[cs], ps1v = P[plane + cs], ps2v = P[2 * plane + cs];
  double su0 = pw0 + pc0, su1 = pw1 + pc1, su2 = pw2 + pc2;  // u pair sums, levels 0,1,2
  double sv0 = ps0v + pc0, sv1 = ps1v + pc1, sv2 = ps2v + pc2;
  double dpn_u = 0.5 * ((g.cf1 * su0 + g.cf2 * su1) + g.cf3 * su2);  // dpn(0)
  double dpn_v = 0.5 * ((g.cf1 * sv0 + g.cf2 * sv1) + g.cf3 * sv2);
  // carried pair sums: cur = level k, prev = k-1, prev2 = k-2
  double su_cur = su0, su_prev = 0.0, su_prev2 = 0.0;
  double sv_cur = sv0, sv_prev = 0.0, sv_prev2 = 0.0;
  double su_next = su1, sv_next = sv1;  // level k+1
  double phc = w.ph_work[cc], phw = w.ph_work[cw], phs = w.ph_work[cs];  // level k
  for (int k = 0; k < nz; ++k) {
    const size_t ok = (size_t)k * plane;
    // dpn(k+1)
    double dpn_u1, dpn_v1;
    if (k + 1 < nz) {
      dpn_u1 = 0.5 * (g.fnm[k + 1] * su_next + g.fnp[k + 1] * su_cur);
      dpn_v1 = 0.5 * (g.fnm[k + 1] * sv_next + g.fnp[k + 1] * sv_cur);
    } else {  // rigid-lid top face: cf1/cf2/cf3 on levels nz-1, nz-2, nz-3
      dpn_u1 = 0.5 * ((g.cf1 * su_cur + g.cf2 * su_prev) + g.cf3 * su_prev2);
      dpn_v1 = 0.5 * ((g.cf1 * sv_cur + g.cf2 * sv_prev) + g.cf3 * sv_prev2);
    }
    const double phc1 = w.ph_work[ok + plane + cc];
    const double altc = w.alt[ok + cc], alc = w.al[ok + cc], pbc = w.pb[ok + cc], phpc = w.php[ok + cc];
    const double pck = P[ok + cc];
    if (do_u) {
      const double phw1 = w.ph_work[ok + plane + cw];
      size_t q = iu(g, k, j, i);
      double u = w.u_work[q] + dts * w.ut[q];
      double mass_x = g.c1h[k] * muu + g.c2h[k];
      double ph_term = (phc1 - phw1) + (phc - phw);
      double p_term = (w.alt[ok + cw] + altc) * (pck - P[ok + cw]);
      double pb_term = (w.al[ok + cw] + alc) * (pbc - w.pb[ok + cw]);
      double dpx = (((msf_u * 0.5) * g.rdx) * mass_x) * ((ph_term + p_term) + pb_term);
      double bracket = g.rdnw[k] * (dpn_u1 - dpn_u) - 0.5 * (g.c1h[k] * muwsum_u);
      dpx = dpx + ((msf_u * g.rdx) * (phpc - w.php[ok + cw])) * bracket;
      u = u - (dts * 1.0) * dpx;
      u = u + g.c1h[k] * mudf_u;
      w.u_work[q] = u;
      phw = phw1;
    }
    if (do_v) {
      const double phs1 = w.ph_work[ok + plane + cs];
      size_t q = iv(g, k, j, i);
      double v = w.v_work[q] + dts * w.vt[q];
      double mass_y = g.c1h[k] * muv + g.c2h[k];
      double ph_term = (phc1 - phs1) + (phc - phs);
      double p_term = (w.alt[ok + cs] + altc) * (pck - P[ok + cs]);
      double pb_term = (w.al[ok + cs] + alc) * (pbc - w.pb[ok + cs]);
      double dpy = (((msf_v * 0.5) * g.rdy) * mass_y) * ((ph_term + p_term) + pb_term);
      double bracket = g.rdnw[k] * (dpn_v1 - dpn_v) - 0.5 * (g.c1h[k] * muwsum_v);
      dpy = dpy + ((msf_v * g.rdy) * (phpc - w.php[ok + cs])) * bracket;
      v = v - (dts * 1.0) * dpy;
      v = v + g.c1h[k] * mudf_v;
      w.v_work[q] = v;
      phs = phs1;
    }
    phc = phc1;
    dpn_u = dpn_u1;
    dpn_v = dpn_v1;
    su_prev2 = su_prev;
    su_prev = su_cur;
    su_cur = su_next;
    sv_prev2 = sv_prev;
    sv_prev = sv_cur;
    sv_cur = sv_next;
    if (k + 2 < nz) {
      const size_t o2 = (size_t)(k + 2) * plane;
      double pc_n = P[o2 + cc];
      su_next = P[o2 + cw] + pc_n;
      sv_next = P[o2 + cs] + pc_n;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused column kernel (one workgroup = TX adjacent columns of one row j,
// KT threads per column):
//   advance_mu_t periodic   dynamics/mu_t_advance.py:139-292
//   advance_w_wrf           core/advance_w.py:293-627 (w_damp :92-132,
//                           Thomas sweeps :558-592, damp_opt=3 :594-617)
//   calc_p_rho_step         core/calc_p_rho.py:142-183 (smdiv = 0.1)
// Level-local work runs in parallel over k (thread kt takes levels kt,
// kt+KT, ...; loads coalesce along i) and is exchanged through LDS; the
// vertical recurrences (the dmdt sum, the omega recurrence, the Thomas
// sweeps) run one thread per column in the reference's order, so every
// value is produced by the same operations as the sequential form.
//
// LDS: six (nz+1) x TX double planes plus three TX rows.
//   A: dvdxi -> explicit w -> solved w      B: omega -> a -> new ph work
//   C: t_2ave(new) -> alpha                 D: phi-equation rhs
//   E: omega_save (rom) -> new coupled theta
//   F: old ph work -> gamma
// ---------------------------------------------------------------------------
template <int TX, int KT>
__global__ void __launch_bounds__(TX * KT) k_column(DGrid g, DNamelist nl, DState s, DWork w, double dts) {
  extern __shared__ double lds[];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int tx = threadIdx.x % TX;
  const int kt = threadIdx.x / TX;
  const int i = blockIdx.x * TX + tx;
  const int j = blockIdx.y;
  const bool valid = i < nx;
  const size_t pl = (size_t)(nz + 1) * TX;
  double* A = lds;
  double* B = A + pl;
  double* C = B + pl;
  double* D = C + pl;
  double* E = D + pl;
  double* F = E + pl;
  double* s_muts = F + pl;
  double* s_muave = s_muts + TX;
  double* s_ph0 = s_muave + TX;
#define L(arr, k) arr[(size_t)(k) * TX + tx]
  const size_t c = valid ? i2(g, j, i) : 0;
  const double msftx = valid ? g.msftx[c] : 1.0;
  const double msfty = valid ? g.msfty[c] : 1.0;
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double mut = valid ? w.mut[c] : 0.0;

  // ---- P1: dvdxi per level (advance_mu_t), omega_save into LDS ----
  double p2_muts = 0.0, p2_mu = 0.0, p2_ww0 = 0.0, p2_wwtop = 0.0;
  if (valid && kt == 0) {  // column values the recurrence thread needs in P2
    p2_muts = w.muts[c];
    p2_mu = w.mu[c];
    p2_ww0 = w.ww[im(g, 0, j, i)];
    p2_wwtop = w.ww[im(g, nz, j, i)];
  }
  if (valid) {
    for (int k = kt; k < nz; k += KT) {
      double c1 = g.c1h[k], c2 = g.c2h[k];
      double v_n = w.v_work[iv(g, k, j + 1, i)] +
                   ((c1 * w.muv[i2v(g, j + 1, i)] + c2) * s.v[iv(g, k, j + 1, i)]) * (1.0 / g.msfvx[i2v(g, j + 1, i)]);
      double v_s = w.v_work[iv(g, k, j, i)] +
                   ((c1 * w.muv[i2v(g, j, i)] + c2) * s.v[iv(g, k, j, i)]) * (1.0 / g.msfvx[i2v(g, j, i)]);
      double u_e = w.u_work[iu(g, k, j, i + 1)] +
                   ((c1 * w.muu[i2u(g, j, i + 1)] + c2) * s.u[iu(g, k, j, i + 1)]) / g.msfuy[i2u(g, j, i + 1)];
      double u_w = w.u_work[iu(g, k, j, i)] +
                   ((c1 * w.muu[i2u(g, j, i)] + c2) * s.u[iu(g, k, j, i)]) / g.msfuy[i2u(g, j, i)];
      L(A, k) = (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
    }
    for (int k = kt; k <= nz; k += KT) L(E, k) = w.rom[im(g, k, j, i)];
  }
  __syncthreads();

  // ---- P2 (one thread per column): dmdt, mass updates, omega recurrence ----
  if (valid && kt == 0) {
    double dmdt = 0.0;
    for (int k = 0; k < nz; ++k) dmdt = (k == 0) ? g.dnw[0] * L(A, 0) : dmdt + g.dnw[k] * L(A, k);
    const double mu_work_old = p2_muts - mut;
    const double mu_save = p2_mu - mu_work_old;
    const double mu_tendency = dmdt + 0.0;
    const double mu_work_new = mu_work_old + dts * mu_tendency;
    w.mu[c] = mu_save + mu_work_new;
    s.mudf[c] = mu_tendency;
    const double muts = mut + mu_work_new;
    w.muts[c] = muts;
    const double muave = 0.5 * (eps_p * mu_work_new + eps_m * mu_work_old);
    w.muave[c] = muave;
    s_muts[tx] = muts;
    s_muave[tx] = muave;
    double wprev = p2_ww0;
    L(B, 0) = wprev - L(E, 0);
    for (int kk = 1; kk < nz; ++kk) {
      int k = kk - 1;
      double incr = (g.dnw[kk - 1] * ((g.c1h[k] * dmdt + L(A, kk - 1)) + g.c1h[k] * 0.0)) / msfty;
      double wk = wprev - incr;
      L(B, kk) = wk - L(E, kk);
      wprev = wk;
    }
    L(B, nz) = p2_wwtop;
  }
  __syncthreads();

  const double muts = valid ? s_muts[tx] : 0.0;
  const double muave = valid ? s_muave[tx] : 0.0;

  // ---- P3: coupled theta update, omega out, new t_2ave ----
  if (valid) {
    int ip = wrapi(i + 1, nx), imn = wrapi(i - 1, nx), jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny);
    auto T1 = [&](int k, int jj, int ii) { return s.theta[im(g, k, jj, ii)] - kT0; };
    for (int k = kt; k < nz; k += KT) {
      double wd_hi = 0.0, wd_lo = 0.0;
      if (k + 1 < nz) {
        int kk = k + 1;
        wd_hi = L(B, kk) * (g.fnm[kk] * T1(kk, j, i) + g.fnp[kk] * T1(kk - 1, j, i));
      }
      if (k >= 1) wd_lo = L(B, k) * (g.fnm[k] * T1(k, j, i) + g.fnp[k] * T1(k - 1, j, i));
      double th = T1(k, j, i);
      double v_flux = w.v_work[iv(g, k, j + 1, i)] * (T1(k, jp, i) + th) - w.v_work[iv(g, k, j, i)] * (th + T1(k, jm, i));
      double u_flux = w.u_work[iu(g, k, j, i + 1)] * (T1(k, j, ip) + th) - w.u_work[iu(g, k, j, i)] * (th + T1(k, j, imn));
      double tendency = msftx * ((0.5 * g.rdy) * v_flux + (0.5 * g.rdx) * u_flux) + g.rdnw[k] * (wd_hi - wd_lo);
      size_t q = im(g, k, j, i);
      double theta_i = w.th_work[q] + ((msfty * dts) * w.tht[q]);
      double th_new = theta_i - ((dts * msfty) * tendency);
      w.th_work[q] = th_new;
      L(E, k) = th_new;
      // t_2ave (advance_w.py:448-458)
      double mass_h = g.c1h[k] * muts + g.c2h[k];
      double half = 0.5 * (eps_p * th_new + eps_m * w.t2ave[q]);
      double ttr = kT0 + th;
      double t2n = (half + (g.c1h[k] * muave) * kT0) / (mass_h * ttr);
      L(C, k) = t2n;
      w.t2ave[q] = t2n;
    }
    for (int f = kt; f <= nz; f += KT) {
      w.ww[im(g, f, j, i)] = L(B, f);
      double phw = w.ph_work[im(g, f, j, i)];
      L(F, f) = phw;
      if (f == 0) s_ph0[tx] = phw;
    }
  }
  __syncthreads();

  // ---- P4: rhs of the phi equation (advance_w.py:460-503) ----
  auto PH1T = [&](int k) { return w.ph0[im(g, k, j, i)] + w.phb[im(g, k, j, i)]; };
  if (valid) {
    const double c05geps = (0.5 * kG) * eps_m;
    for (int f = kt; f <= nz; f += KT) {
      if (f == 0) {
        L(D, 0) = 0.0;
        continue;
      }
      double r = dts * (w.ph_tend[im(g, f, j, i)] + c05geps * w.w_work[im(g, f, j, i)]);
      if (f < nz) {
        double wd_f1 = ((0.5 * (L(B, f + 1) + L(B, f))) * g.rdnw[f]) * (PH1T(f + 1) - PH1T(f));
        double wd_f = ((0.5 * (L(B, f) + L(B, f - 1))) * g.rdnw[f - 1]) * (PH1T(f) - PH1T(f - 1));
        double adv = dts * (g.fnm[f] * wd_f1 + g.fnp[f] * wd_f);
        r = r + (-adv);
      }
      double mass_f = g.c1f[f] * mut + g.c2f[f];
      r = L(F, f) + (msfty * r) / mass_f;
      if (f == nz && nl.top_lid) r = 0.0;
      L(D, f) = r;
    }
  }
  __syncthreads();

  // ---- P5: explicit w (advance_w.py:505-556) and the terrain surface w ----
  if (valid) {
    const double msft_inv = 1.0 / msfty;
    const double c05dtg = (0.5 * dts) * kG;
    for (int f = kt; f <= nz; f += KT) {
      if (f == 0) {
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
        L(A, 0) = ((msfty * 0.5) * g.rdy) * (dy_n * v_cf_n + dy_s * v_cf_s) +
                  ((msftx * 0.5) * g.rdx) * (dx_e * u_cf_e + dx_w * u_cf_w);
        continue;
      }
      if (f == nz) {
        L(A, nz) = 0.0;  // rigid lid
        continue;
      }
      size_t qf = im(g, f, j, i);
      double wold = w.w_work[qf];
      double rw = w.rw_tend[qf];
      if (nl.w_damping == 1) {
        double mass_f = g.c1f[f] * mut + g.c2f[f];
        double safe_mf = safe_pos(mass_f, 1.0e-12);
        double vert_cfl = fabs(((L(B, f) / safe_mf) * g.rdnw[f]) * dts);
        double damp = (((-sgn(wold)) * kWAlpha) * (vert_cfl - kWBeta)) * mass_f;
        rw = rw + (vert_cfl > kWBeta ? damp : 0.0);
      }
      double wn = wold + dts * rw;
      size_t qm = im(g, f - 1, j, i);
      double mh_u = safe_pos(g.c1h[f] * mut + g.c2h[f], 1.0e-12);
      double mh_l = safe_pos(g.c1h[f - 1] * mut + g.c2h[f - 1], 1.0e-12);
      double c2a_u = w.c2a[qf], c2a_l = w.c2a[qm];
      double coef_u = (c2a_u * g.rdnw[f]) / mh_u;
      double coef_l = (c2a_l * g.rdnw[f - 1]) / mh_l;
      double phk1 = L(F, f + 1), phk = L(F, f), phkm = L(F, f - 1);
      double tA_u = coef_u * (eps_p * (L(D, f + 1) - L(D, f)) + eps_m * (phk1 - phk));
      double tA_l = coef_l * (eps_p * (L(D, f) - L(D, f - 1)) + eps_m * (phk - phkm));
      double termA = ((msft_inv * cqw_at(f, nz)) * (c05dtg * g.rdn[f])) * (tA_u - tA_l);
      double bu = (c2a_u * w.alt[qf]) * L(C, f);
      double bl = (c2a_l * w.alt[qm]) * L(C, f - 1);
      double termB = ((dts * kG) * msft_inv) * (g.rdn[f] * (bu - bl) - (g.c1f[f] * muave));
      L(A, f) = wn + (termA + termB);
    }
  }
  __syncthreads();

  // ---- stage the tridiagonal coefficients (B, C, F are free now) ----
  if (valid) {
    for (int f = kt; f <= nz; f += KT) {
      size_t qf = im(g, f, j, i);
      L(B, f) = w.ca[qf];
      L(C, f) = w.calpha[qf];
      L(F, f) = w.cgamma[qf];
    }
  }
  __syncthreads();

  // ---- P6 (one thread per column): Thomas forward / back (advance_w.py:558-592) ----
  if (valid && kt == 0) {
    double prev = L(A, 0);
    for (int f = 1; f <= nz; ++f) {
      double wf = (L(A, f) - L(B, f) * prev) * L(C, f);
      L(A, f) = wf;
      prev = wf;
    }
    double next = L(A, nz);
    for (int f = nz - 1; f >= 1; --f) {
      double wv = L(A, f) - L(F, f) * next;
      L(A, f) = wv;
      next = wv;
    }
  }
  __syncthreads();

  // ---- P7: damp_opt=3 Rayleigh damping and the geopotential finish ----
  if (valid) {
    const bool damp = (nl.damp_opt == 3 && nl.dampcoef > 0.0);
    const double dampmag = dts * nl.dampcoef, hdepth = nl.zdamp;
    const double htop = PH1T(nz) / kG;
    const double hbot = htop - hdepth;
    for (int f = kt; f <= nz; f += KT) {
      size_t qf = im(g, f, j, i);
      if (f == 0) {
        w.w_work[qf] = L(A, 0);
        L(B, 0) = s_ph0[tx];
        continue;
      }
      double wv = L(A, f);
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
      double phn = L(D, f) + (((((msfty * 0.5) * dts) * kG) * eps_p) * wd) / mfs;
      w.ph_work[qf] = phn;
      w.w_work[qf] = wd;
      L(B, f) = phn;
    }
  }
  __syncthreads();

  // ---- P8: calc_p_rho_step ----
  if (valid) {
    const double mu_work = muts - mut;
    for (int k = kt; k < nz; k += KT) {
      size_t q = im(g, k, j, i);
      double mass_h = g.c1h[k] * muts + g.c2h[k];
      double safe_mass = safe_pos(mass_h, 1.0e-12);
      double mu_term = g.c1h[k] * mu_work;
      double alt = w.alt[q];
      double al = -(alt * mu_term + g.rdnw[k] * (L(B, k + 1) - L(B, k))) / safe_mass;
      double theta_1 = s.theta[q] - kT0;
      double ttr = kT0 + theta_1;
      double safe_t = safe_pos(ttr, 1.0e-6);
      double p = w.c2a[q] * (((alt * (L(E, k) - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
      double ptmp = p;
      p = p + kSmdiv * (p - w.pm1[q]);
      w.al[q] = al;
      w.p[q] = p;
      w.pm1[q] = ptmp;
    }
  }
#undef L
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

template <int TX, int KT>
hipError_t launch_column(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, double dts,
                         hipStream_t st, Prof* prof) {
  dim3 colgrid((g.nx + TX - 1) / TX, g.ny);
  size_t lds = (6 * (size_t)(g.nz + 1) * TX + 3 * TX) * sizeof(double);
  WD_TIMED(prof, KID_COLUMN, st, (k_column<TX, KT><<<colgrid, TX * KT, lds, st>>>(g, nl, s, w, dts)));
  return hipGetLastError();
}

}  // namespace

// Column-tile shapes (TX columns x KT threads per column) compiled in.
const ColumnTile kColumnTiles[] = {{16, 16}, {32, 8}, {16, 8}, {32, 4}, {64, 4}, {8, 16}};
const int kNumColumnTiles = sizeof(kColumnTiles) / sizeof(kColumnTiles[0]);

namespace {

hipError_t launch_column_tile(int tile, const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              double dts, hipStream_t st, Prof* prof) {
  switch (tile) {
    case 1: return launch_column<32, 8>(g, nl, s, w, dts, st, prof);
    case 2: return launch_column<16, 8>(g, nl, s, w, dts, st, prof);
    case 3: return launch_column<32, 4>(g, nl, s, w, dts, st, prof);
    case 4: return launch_column<64, 4>(g, nl, s, w, dts, st, prof);
    case 5: return launch_column<8, 16>(g, nl, s, w, dts, st, prof);
    default: return launch_column<16, 16>(g, nl, s, w, dts, st, prof);
  }
}

}  // namespace

hipError_t launch_acoustic_substep(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                   double dts, int column_tile, hipStream_t st, Prof* prof) {
  constexpr int BX = 64, BY = 4;
  dim3 uvgrid((g.nx + 1 + BX - 1) / BX, (g.ny + 1 + BY - 1) / BY);
  WD_TIMED(prof, KID_UV, st, (k_uv<BX, BY><<<uvgrid, dim3(BX, BY), 0, st>>>(g, s, w, dts)));
  return launch_column_tile(column_tile, g, nl, s, w, dts, st, prof);
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
      "k_uv", "k_column", "k_fin_w", "k_fin_uv", "k_fin_m", "k_fin_mu"};
  return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

}  // namespace wd
