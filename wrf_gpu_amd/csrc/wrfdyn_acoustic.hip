// Acoustic small step (advance_uv, then the fused column kernel
// advance_mu_t + advance_w + calc_p_rho) and the stage finish
// (small_step_finish + grid%p refresh).
//
// Each output value is computed with the reference's expression order; build
// with -ffp-contract=off.
#include "wrfdyn_kernels.h"

namespace wd {

namespace {

// Value of lane - 1 / lane + 1 in the wavefront; the lane with no source keeps `old`.
__device__ __forceinline__ double wave_shr1(double old, double src) { return dpp_move<kDppWaveShr1>(old, src); }
__device__ __forceinline__ double wave_shl1(double old, double src) { return dpp_move<kDppWaveShl1>(old, src); }

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
// i < nx; the mass column (j, i) is shared by both stencils.  A block is 64
// faces by BY rows, one wave per row: the west column comes from the lane
// below (DPP) and the south column from the wave below (LDS).  The dpn pair
// sums of the work pressure are carried up the column (interior dpn(k) uses
// levels k and k-1; the bottom and rigid-lid top faces use three levels).
// ---------------------------------------------------------------------------
// WET: the state carries moisture (cqu/cqv scale the pressure gradient); dry launches fold them to 1
template <int BX, int BY, class TT, bool WET>
__device__ __forceinline__ void b_uv(const Blk& B, double* lds, DGrid g, DState s, DWork w,
                                     double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // waves overlap by one lane: lanes 1..BX-1 own faces, lane 0 holds the
  // column west of the first owned face
  const unsigned nbx = (nx + 1 + BX - 2) / (BX - 1);
  const unsigned t = tile_id(g, B);
  const int i = (int)(t % nbx) * (BX - 1) + (int)B.tx - 1;
  const int j = (int)(t / nbx) * BY + B.ty;
  // threads past the face grid, and lane 0, stay resident for the per-level
  // barrier; their mass-column indices are clamped and their face work is
  // switched off
  const bool active = B.tx >= 1 && i <= nx && j <= ny;
  const bool do_u = active && j < ny;
  const bool do_v = active && i < nx;
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
  // Per-level loads, double-buffered: level k+1's loads are issued before
  // level k is computed and stored, so one level of loads is always in flight.
  struct Lv {
    double phc1, phw1, phs1;                  // ph work at level k+1
    double altc, alc, pbc, phpc, pc;          // centre column, level k
    double altw, alw, pbw, phpw, pw, uw, ut;  // west column and the u face
    double alts, als, pbs, phps, ps, vw, vt;  // south column and the v face
    double p2c, p2w, p2s;                     // p at level k+2 (pair sums)
  };
  const unsigned pl8 = (unsigned)plane * 8u;
  const unsigned cc8 = (unsigned)cc * 8u, cs8 = (unsigned)cs * 8u;
  // a wave is one row of BX faces (the block is BX wide, one wavefront)
  static_assert(BX == 64, "k_uv maps one wavefront to one row of faces");
  // row j - 1 is in this block unless this is its first row or j = 0 (whose
  // south column is clamped to row 0 itself)
  const bool row_below_outside = B.ty == 0 || j == 0;
  // face offsets (used only where do_u / do_v hold, i.e. for in-range faces)
  const int ifc = clampi(i, 0, nx), jfc = clampi(j, 0, ny);
  const unsigned uq8 = (unsigned)iu(g, 0, jfc, ifc) * 8u, upl8 = (unsigned)((size_t)ny * (nx + 1)) * 8u;
  const unsigned vq8 = (unsigned)iv(g, 0, jfc, ifc) * 8u, vpl8 = (unsigned)((size_t)(ny + 1) * nx) * 8u;
  auto load = [&](int k, Lv& L) {
    const unsigned ok = (unsigned)k * pl8, ok1 = ok + pl8;
    L.phc1 = ldo(w.ph_work, ok1 + cc8);
    L.altc = ldo(w.alt, ok + cc8);
    L.alc = ldo(w.al, ok + cc8);
    L.pbc = ldo(w.pb, ok + cc8);
    L.phpc = ldo(w.php, ok + cc8);
    L.pc = ldo(P, ok + cc8);
    if (do_u) {
      const unsigned q = uq8 + (unsigned)k * upl8;
      L.uw = ldo(w.u_work, q);
      L.ut = ldt<TT>(w.ut, q);
    }
    if (do_v) {
      const unsigned q = vq8 + (unsigned)k * vpl8;
      if (row_below_outside) {  // rows 1.. of the block take the south column from LDS
        L.phs1 = ldo(w.ph_work, ok1 + cs8);
        L.alts = ldo(w.alt, ok + cs8);
        L.als = ldo(w.al, ok + cs8);
        L.pbs = ldo(w.pb, ok + cs8);
        L.phps = ldo(w.php, ok + cs8);
        L.ps = ldo(P, ok + cs8);
      }
      L.vw = ldo(w.v_work, q);
      L.vt = ldt<TT>(w.vt, q);
    }
    if (k + 2 < nz) {
      const unsigned o2 = ok + 2 * pl8;
      L.p2c = ldo(P, o2 + cc8);
      if (do_v && row_below_outside) L.p2s = ldo(P, o2 + cs8);
    }
  };
  // The south column of row j is the centre column of row j - 1, which the
  // wave below holds (rows j and j - 1 of one block): each level's centre
  // values go through LDS, double-buffered, with one barrier per level.
  double (*sh)[BY][7][BX] = reinterpret_cast<double (*)[BY][7][BX]>(lds);
  auto south = [&](int k, Lv& L) {
    double* mine = &sh[k & 1][B.ty][0][B.tx];
    mine[0 * BX] = L.phc1;
    mine[1 * BX] = L.altc;
    mine[2 * BX] = L.alc;
    mine[3 * BX] = L.pbc;
    mine[4 * BX] = L.phpc;
    mine[5 * BX] = L.pc;
    mine[6 * BX] = L.p2c;
    __syncthreads();
    if (!row_below_outside) {
      const double* below = &sh[k & 1][B.ty - 1][0][B.tx];
      L.phs1 = below[0 * BX];
      L.alts = below[1 * BX];
      L.als = below[2 * BX];
      L.pbs = below[3 * BX];
      L.phps = below[4 * BX];
      L.ps = below[5 * BX];
      L.p2s = below[6 * BX];
    }
  };
  // The west column of face i is the centre column of face i-1, which the
  // next lane down holds: each west value is that lane's centre value moved
  // up one lane by a DPP wave shift (no memory access).  Lane 0 owns no face.
  auto west = [&](Lv& L) {
    L.phw1 = wave_shr1(L.phw1, L.phc1);
    L.altw = wave_shr1(L.altw, L.altc);
    L.alw = wave_shr1(L.alw, L.alc);
    L.pbw = wave_shr1(L.pbw, L.pbc);
    L.phpw = wave_shr1(L.phpw, L.phpc);
    L.pw = wave_shr1(L.pw, L.pc);
    L.p2w = wave_shr1(L.p2w, L.p2c);
  };
  Lv cur{};
  load(0, cur);
  for (int k = 0; k < nz; ++k) {
    Lv nxt{};
    if (k + 1 < nz) load(k + 1, nxt);
    west(cur);
    south(k, cur);
    // dpn(k+1)
    double dpn_u1, dpn_v1;
    if (k + 1 < nz) {
      dpn_u1 = 0.5 * (g.fnm[k + 1] * su_next + g.fnp[k + 1] * su_cur);
      dpn_v1 = 0.5 * (g.fnm[k + 1] * sv_next + g.fnp[k + 1] * sv_cur);
    } else {  // rigid-lid top face: cf1/cf2/cf3 on levels nz-1, nz-2, nz-3
      dpn_u1 = 0.5 * ((g.cf1 * su_cur + g.cf2 * su_prev) + g.cf3 * su_prev2);
      dpn_v1 = 0.5 * ((g.cf1 * sv_cur + g.cf2 * sv_prev) + g.cf3 * sv_prev2);
    }
    const double phc1 = cur.phc1;
    const double altc = cur.altc, alc = cur.alc, pbc = cur.pbc, phpc = cur.phpc;
    const double pck = cur.pc;
    if (do_u) {
      const double phw1 = cur.phw1;
      const unsigned q = uq8 + (unsigned)k * upl8;
      double u = cur.uw + dts * cur.ut;
      double mass_x = g.c1h[k] * muu + g.c2h[k];
      double ph_term = (phc1 - phw1) + (phc - phw);
      double p_term = (cur.altw + altc) * (pck - cur.pw);
      double pb_term = (cur.alw + alc) * (pbc - cur.pbw);
      double dpx = (((msf_u * 0.5) * g.rdx) * mass_x) * ((ph_term + p_term) + pb_term);
      double bracket = g.rdnw[k] * (dpn_u1 - dpn_u) - 0.5 * (g.c1h[k] * muwsum_u);
      dpx = dpx + ((msf_u * g.rdx) * (phpc - cur.phpw)) * bracket;
      u = u - (dts * (WET ? ldo(w.cqu, q) : 1.0)) * dpx;
      u = u + g.c1h[k] * mudf_u;
      sto(w.u_work, q, u);
      phw = phw1;
    }
    if (do_v) {
      const double phs1 = cur.phs1;
      const unsigned q = vq8 + (unsigned)k * vpl8;
      double v = cur.vw + dts * cur.vt;
      double mass_y = g.c1h[k] * muv + g.c2h[k];
      double ph_term = (phc1 - phs1) + (phc - phs);
      double p_term = (cur.alts + altc) * (pck - cur.ps);
      double pb_term = (cur.als + alc) * (pbc - cur.pbs);
      double dpy = (((msf_v * 0.5) * g.rdy) * mass_y) * ((ph_term + p_term) + pb_term);
      double bracket = g.rdnw[k] * (dpn_v1 - dpn_v) - 0.5 * (g.c1h[k] * muwsum_v);
      dpy = dpy + ((msf_v * g.rdy) * (phpc - cur.phps)) * bracket;
      v = v - (dts * (WET ? ldo(w.cqv, q) : 1.0)) * dpy;
      v = v + g.c1h[k] * mudf_v;
      sto(w.v_work, q, v);
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
      double pc_n = cur.p2c;
      su_next = cur.p2w + pc_n;
      sv_next = cur.p2s + pc_n;
    }
    cur = nxt;
  }
}

template <int BX, int BY, class TT, bool WET>
__global__ void __launch_bounds__(BX * BY) k_uv(DGrid g, DState s, DWork w, double dts) {
  __shared__ double sh[2][BY][7][BX];
  b_uv<BX, BY, TT, WET>(hw_blk(), &sh[0][0][0][0], g, s, w, dts);
}

// ---------------------------------------------------------------------------
// Fused column kernel (one workgroup = TX adjacent columns of one row j,
// KT threads per column, each thread owning the R levels kt, kt+KT, ...):
//   advance_mu_t periodic   dynamics/mu_t_advance.py:139-292
//   advance_w_wrf           core/advance_w.py:293-627 (w_damp :92-132,
//                           Thomas sweeps :558-592, damp_opt=3 :594-617)
//   calc_p_rho_step         core/calc_p_rho.py:142-183 (smdiv = 0.1)
// Level-local work runs in parallel over k and is exchanged through LDS; the
// vertical recurrences (the dmdt sum, the omega recurrence, the Thomas
// sweeps) run one thread per column in the reference's order, so every value
// is produced by the same operations as the sequential form.
//
// Every phase first issues all of its global loads for the thread's R levels
// into registers and only then computes, so a wave keeps R levels of loads in
// flight instead of one; values a later phase needs again (theta, c2a, alt,
// ph_1 total, the old w) stay in registers instead of being re-read.
//
// Up to four levels per thread the kernel is held to 168 VGPRs (three waves
// per SIMD) without spilling.  The serial sweeps read LDS and the metrics
// eight levels ahead of their arithmetic chain.
//
// LDS: six (nz+1) x TX double planes plus three TX rows.
//   A: dvdxi -> explicit w -> solved w      B: omega -> a -> new ph work
//   C: t_2ave(new) -> alpha                 D: phi-equation rhs
//   E: omega_save (rom) -> new coupled theta
//   F: old ph work -> gamma
// ---------------------------------------------------------------------------
template <int TX, int KT, int R>
__device__ __forceinline__ void b_column(const int tx, const int kt, const unsigned t, double* lds, DGrid g,
                                         DNamelist nl, DState s, DWork w, double dts, int first) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned nbx = (nx + TX - 1) / TX;
  const int i = (int)(t % nbx) * TX + tx;
  const int j = (int)(t / nbx);
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
  // Flat offsets: mass/w level k of column (j, i) is c + k * plane.
  const unsigned plane = (unsigned)((size_t)ny * nx) * 8u;  // byte strides (ldo/sto)
  const unsigned uplane = (unsigned)((size_t)ny * (nx + 1)) * 8u, vplane = (unsigned)((size_t)(ny + 1) * nx) * 8u;
  const size_t c = valid ? i2(g, j, i) : 0;
  const unsigned c8 = (unsigned)c * 8u;
  const double msftx = valid ? g.msftx[c] : 1.0;
  const double msfty = valid ? g.msfty[c] : 1.0;
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double mut = valid ? w.mut[c] : 0.0;
  auto lev = [&](int r) { return kt + r * KT; };

  // ---- P1: dvdxi per level (advance_mu_t), omega_save into LDS ----
  double p2_muts = 0.0, p2_mu = 0.0, p2_ww0 = 0.0, p2_wwtop = 0.0;
  if (valid && kt == 0) {  // column values the recurrence thread needs in P2
    p2_muts = w.muts[c];
    p2_mu = w.mu[c];
    p2_ww0 = w.ww[c];
    p2_wwtop = ldo(w.ww, c8 + (unsigned)nz * plane);
  }
  {
    const size_t fvn = valid ? i2v(g, j + 1, i) : 0, fvs = valid ? i2v(g, j, i) : 0;
    const size_t fue = valid ? i2u(g, j, i + 1) : 0, fuw = valid ? i2u(g, j, i) : 0;
    const unsigned vn8 = (unsigned)fvn * 8u, vs8 = (unsigned)fvs * 8u, ue8 = (unsigned)fue * 8u, uw8 = (unsigned)fuw * 8u;
    double muv_n = 0.0, muv_s = 0.0, rmvx_n = 0.0, rmvx_s = 0.0;
    double muu_e = 0.0, muu_w = 0.0, msfuy_e = 1.0, msfuy_w = 1.0;
    if (valid) {
      muv_n = w.muv[fvn];
      muv_s = w.muv[fvs];
      rmvx_n = 1.0 / g.msfvx[fvn];
      rmvx_s = 1.0 / g.msfvx[fvs];
      muu_e = w.muu[fue];
      muu_w = w.muu[fuw];
      msfuy_e = g.msfuy[fue];
      msfuy_w = g.msfuy[fuw];
    }
    double vwn[R], vws[R], vn[R], vs[R], uwe[R], uww[R], ue[R], uw[R], romr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      if (valid && k < nz) {
        const unsigned kv = (unsigned)k * vplane, ku = (unsigned)k * uplane;
        vwn[r] = ldo(w.v_work, vn8 + kv);
        vws[r] = ldo(w.v_work, vs8 + kv);
        vn[r] = ldo(s.v, vn8 + kv);
        vs[r] = ldo(s.v, vs8 + kv);
        uwe[r] = ldo(w.u_work, ue8 + ku);
        uww[r] = ldo(w.u_work, uw8 + ku);
        ue[r] = ldo(s.u, ue8 + ku);
        uw[r] = ldo(s.u, uw8 + ku);
      }
      if (valid && k <= nz) romr[r] = ldo(w.rom, c8 + (unsigned)k * plane);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      if (valid && k < nz) {
        const double c1 = g.c1h[k], c2 = g.c2h[k];
        double v_n = vwn[r] + ((c1 * muv_n + c2) * vn[r]) * rmvx_n;
        double v_s = vws[r] + ((c1 * muv_s + c2) * vs[r]) * rmvx_s;
        double u_e = uwe[r] + ((c1 * muu_e + c2) * ue[r]) / msfuy_e;
        double u_w = uww[r] + ((c1 * muu_w + c2) * uw[r]) / msfuy_w;
        L(A, k) = (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
      }
      if (valid && k <= nz) L(E, k) = romr[r];
    }
  }
  __syncthreads();

  // ---- P2 (one thread per column): dmdt, mass updates, omega recurrence ----
  if (valid && kt == 0) {
    // LDS reads are batched eight levels ahead of the serial arithmetic
    double dmdt = 0.0;
    for (int k0 = 0; k0 < nz; k0 += 8) {
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = k0 + u < nz ? L(A, k0 + u) : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        if (k < nz) dmdt = (k == 0) ? g.dnw[0] * a[u] : dmdt + g.dnw[k] * a[u];
      }
    }
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
    const double ww0 = wprev - L(E, 0);
    L(B, 0) = ww0;
    w.ww[c] = ww0;  // carried to the next substep; the interior is rebuilt each substep
    for (int k0 = 1; k0 < nz; k0 += 8) {
      double a[8], e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k0 + u;
        a[u] = kk < nz ? L(A, kk - 1) : 0.0;
        e[u] = kk < nz ? L(E, kk) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k0 + u;
        if (kk < nz) {
          const int k = kk - 1;
          double incr = (g.dnw[kk - 1] * ((g.c1h[k] * dmdt + a[u]) + g.c1h[k] * 0.0)) / msfty;
          double wk = wprev - incr;
          L(B, kk) = wk - e[u];
          wprev = wk;
        }
      }
    }
    L(B, nz) = p2_wwtop;
  }
  __syncthreads();

  const double muts = valid ? s_muts[tx] : 0.0;
  const double muave = valid ? s_muave[tx] : 0.0;

  // ---- P3: coupled theta update, new t_2ave, old ph work into LDS ----
  double thc[R];  // theta at the thread's levels, reused by calc_p_rho
  {
    const int ip = valid ? wrapi(i + 1, nx) : 0, imn = valid ? wrapi(i - 1, nx) : 0;
    const int jp = wrapi(j + 1, ny), jm = wrapi(j - 1, ny);
    const unsigned cjp = (unsigned)i2(g, jp, i) * 8u, cjm = (unsigned)i2(g, jm, i) * 8u;
    const unsigned cip = (unsigned)i2(g, j, ip) * 8u, cim = (unsigned)i2(g, j, imn) * 8u;
    const unsigned vn8 = valid ? (unsigned)i2v(g, j + 1, i) * 8u : 0u, vs8 = valid ? (unsigned)i2v(g, j, i) * 8u : 0u;
    const unsigned ue8 = valid ? (unsigned)i2u(g, j, i + 1) * 8u : 0u, uw8 = valid ? (unsigned)i2u(g, j, i) * 8u : 0u;
    double thm[R], thp[R], thn[R], ths[R], the[R], thw[R], vwn[R], vws[R], uwe[R], uww[R];
    double thwk[R], tht[R], t2o[R], phw[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      if (valid && k < nz) {
        const unsigned kp = (unsigned)k * plane, q = c8 + kp;
        const unsigned kv = (unsigned)k * vplane, ku = (unsigned)k * uplane;
        thc[r] = ldo(s.theta, q);
        thm[r] = k >= 1 ? ldo(s.theta, q - plane) : 0.0;
        thp[r] = k + 1 < nz ? ldo(s.theta, q + plane) : 0.0;
        thn[r] = ldo(s.theta, cjp + kp);
        ths[r] = ldo(s.theta, cjm + kp);
        the[r] = ldo(s.theta, cip + kp);
        thw[r] = ldo(s.theta, cim + kp);
        vwn[r] = ldo(w.v_work, vn8 + kv);
        vws[r] = ldo(w.v_work, vs8 + kv);
        uwe[r] = ldo(w.u_work, ue8 + ku);
        uww[r] = ldo(w.u_work, uw8 + ku);
        thwk[r] = ldo(w.th_work, q);
        tht[r] = ldo(w.tht, q);
        t2o[r] = first ? 0.0 : ldo(w.t2ave, q);  // t_2ave starts each stage at 0
      }
      if (valid && k <= nz) phw[r] = ldo(w.ph_work, c8 + (unsigned)k * plane);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      if (valid && k < nz) {
        const double th = thc[r] - kT0;
        double wd_hi = 0.0, wd_lo = 0.0;
        if (k + 1 < nz) {
          const int kk = k + 1;
          wd_hi = L(B, kk) * (g.fnm[kk] * (thp[r] - kT0) + g.fnp[kk] * th);
        }
        if (k >= 1) wd_lo = L(B, k) * (g.fnm[k] * th + g.fnp[k] * (thm[r] - kT0));
        double v_flux = vwn[r] * ((thn[r] - kT0) + th) - vws[r] * (th + (ths[r] - kT0));
        double u_flux = uwe[r] * ((the[r] - kT0) + th) - uww[r] * (th + (thw[r] - kT0));
        double tendency = msftx * ((0.5 * g.rdy) * v_flux + (0.5 * g.rdx) * u_flux) + g.rdnw[k] * (wd_hi - wd_lo);
        const unsigned q = c8 + (unsigned)k * plane;
        double theta_i = thwk[r] + ((msfty * dts) * tht[r]);
        double th_new = theta_i - ((dts * msfty) * tendency);
        sto(w.th_work, q, th_new);
        L(E, k) = th_new;
        // t_2ave (advance_w.py:448-458)
        double mass_h = g.c1h[k] * muts + g.c2h[k];
        double half = 0.5 * (eps_p * th_new + eps_m * t2o[r]);
        double ttr = kT0 + th;
        double t2n = (half + (g.c1h[k] * muave) * kT0) / (mass_h * ttr);
        L(C, k) = t2n;
        sto(w.t2ave, q, t2n);
      }
      if (valid && k <= nz) {
        L(F, k) = phw[r];
        if (k == 0) s_ph0[tx] = phw[r];
      }
    }
  }
  __syncthreads();

  // ---- P4: rhs of the phi equation (advance_w.py:460-503) ----
  double ph1c[R], wold[R];  // ph_1 total and the old w at the thread's levels
  {
    double pht[R], php1[R], phm1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int f = lev(r);
      if (valid && f <= nz) {
        const unsigned q = c8 + (unsigned)f * plane;
        ph1c[r] = ldo(w.ph1t, q);
        wold[r] = ldo(w.w_work, q);
        pht[r] = ldo(w.ph_tend, q);
        php1[r] = f < nz ? ldo(w.ph1t, q + plane) : 0.0;
        phm1[r] = f >= 1 ? ldo(w.ph1t, q - plane) : 0.0;
      }
    }
    const double c05geps = (0.5 * kG) * eps_m;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int f = lev(r);
      if (!(valid && f <= nz)) continue;
      if (f == 0) {
        L(D, 0) = 0.0;
        continue;
      }
      double rr = dts * (pht[r] + c05geps * wold[r]);
      if (f < nz) {
        double wd_f1 = ((0.5 * (L(B, f + 1) + L(B, f))) * g.rdnw[f]) * (php1[r] - ph1c[r]);
        double wd_f = ((0.5 * (L(B, f) + L(B, f - 1))) * g.rdnw[f - 1]) * (ph1c[r] - phm1[r]);
        double adv = dts * (g.fnm[f] * wd_f1 + g.fnp[f] * wd_f);
        rr = rr + (-adv);
      }
      double mass_f = g.c1f[f] * mut + g.c2f[f];
      rr = L(F, f) + (msfty * rr) / mass_f;
      if (f == nz && nl.top_lid) rr = 0.0;
      L(D, f) = rr;
    }
  }
  __syncthreads();

  // ---- P5: explicit w (advance_w.py:505-556) and the terrain surface w;
  //      the tridiagonal coefficients are loaded here and staged after ----
  double c2ar[R], altr[R];  // c2a and alt at the thread's mass levels, reused by calc_p_rho
  double car[R], calr[R], cgar[R];
  {
    double rwt[R], c2am[R], altm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int f = lev(r);
      if (valid && f <= nz) {
        const unsigned q = c8 + (unsigned)f * plane;
        car[r] = ldo(w.ca, q);
        calr[r] = ldo(w.calpha, q);
        cgar[r] = ldo(w.cgamma, q);
        if (f < nz) {
          c2ar[r] = ldo(w.c2a, q);
          altr[r] = ldo(w.alt, q);
        }
        if (f >= 1 && f < nz) {
          rwt[r] = ldo(w.rw_tend, q);
          c2am[r] = ldo(w.c2a, q - plane);
          altm[r] = ldo(w.alt, q - plane);
        }
      }
    }
    const double msft_inv = 1.0 / msfty;
    const double c05dtg = (0.5 * dts) * kG;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int f = lev(r);
      if (!(valid && f <= nz)) continue;
      if (f == 0) {
        const int jn = clampi(j + 1, 0, ny - 1), js = clampi(j - 1, 0, ny - 1);
        const int ie = clampi(i + 1, 0, nx - 1), iw = clampi(i - 1, 0, nx - 1);
        auto PHB0 = [&](int jj, int ii) {
          const size_t q = i2(g, jj, ii);
          return s.ph_total[q] - s.ph_pert[q];
        };
        double ht = PHB0(j, i) / kG;
        double ht_n = PHB0(jn, i) / kG, ht_s = PHB0(js, i) / kG;
        double ht_e = PHB0(j, ie) / kG, ht_w = PHB0(j, iw) / kG;
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
      double rw = rwt[r];
      if (nl.w_damping == 1) {
        double mass_f = g.c1f[f] * mut + g.c2f[f];
        double safe_mf = safe_pos(mass_f, 1.0e-12);
        double vert_cfl = fabs(((L(B, f) / safe_mf) * g.rdnw[f]) * dts);
        double damp = (((-sgn(wold[r])) * kWAlpha) * (vert_cfl - kWBeta)) * mass_f;
        rw = rw + (vert_cfl > kWBeta ? damp : 0.0);
      }
      double wn = wold[r] + dts * rw;
      double mh_u = safe_pos(g.c1h[f] * mut + g.c2h[f], 1.0e-12);
      double mh_l = safe_pos(g.c1h[f - 1] * mut + g.c2h[f - 1], 1.0e-12);
      double c2a_u = c2ar[r], c2a_l = c2am[r];
      double coef_u = (c2a_u * g.rdnw[f]) / mh_u;
      double coef_l = (c2a_l * g.rdnw[f - 1]) / mh_l;
      double phk1 = L(F, f + 1), phk = L(F, f), phkm = L(F, f - 1);
      double tA_u = coef_u * (eps_p * (L(D, f + 1) - L(D, f)) + eps_m * (phk1 - phk));
      double tA_l = coef_l * (eps_p * (L(D, f) - L(D, f - 1)) + eps_m * (phk - phkm));
      double termA = ((msft_inv * cqw_at(f, nz)) * (c05dtg * g.rdn[f])) * (tA_u - tA_l);
      double bu = (c2a_u * altr[r]) * L(C, f);
      double bl = (c2a_l * altm[r]) * L(C, f - 1);
      double termB = ((dts * kG) * msft_inv) * (g.rdn[f] * (bu - bl) - (g.c1f[f] * muave));
      L(A, f) = wn + (termA + termB);
    }
  }
  __syncthreads();

  // ---- stage the tridiagonal coefficients (B, C, F are free now) ----
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = lev(r);
    if (valid && f <= nz) {
      L(B, f) = car[r];
      L(C, f) = calr[r];
      L(F, f) = cgar[r];
    }
  }
  __syncthreads();

  // ---- P6 (one thread per column): Thomas forward / back (advance_w.py:558-592) ----
  if (valid && kt == 0) {
    double prev = L(A, 0);
    for (int f0 = 1; f0 <= nz; f0 += 8) {
      double a[8], b[8], cc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 + u;
        a[u] = f <= nz ? L(A, f) : 0.0;
        b[u] = f <= nz ? L(B, f) : 0.0;
        cc[u] = f <= nz ? L(C, f) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 + u;
        if (f <= nz) {
          double wf = (a[u] - b[u] * prev) * cc[u];
          L(A, f) = wf;
          prev = wf;
        }
      }
    }
    double next = L(A, nz);
    for (int f0 = nz - 1; f0 >= 1; f0 -= 8) {
      double a[8], ff[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 - u;
        a[u] = f >= 1 ? L(A, f) : 0.0;
        ff[u] = f >= 1 ? L(F, f) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int f = f0 - u;
        if (f >= 1) {
          double wv = a[u] - ff[u] * next;
          L(A, f) = wv;
          next = wv;
        }
      }
    }
  }
  __syncthreads();

  // ---- P7: damp_opt=3 Rayleigh damping and the geopotential finish ----
  {
    const bool damp = (nl.damp_opt == 3 && nl.dampcoef > 0.0);
    const double dampmag = dts * nl.dampcoef, hdepth = nl.zdamp;
    // ph_1 total at the lid (thread-independent column value; the owner of
    // level nz holds it, everyone reads it from global once)
    const double htop = valid ? ldo(w.ph1t, c8 + (unsigned)nz * plane) / kG : 0.0;
    const double hbot = htop - hdepth;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int f = lev(r);
      if (!(valid && f <= nz)) continue;
      const unsigned qf = c8 + (unsigned)f * plane;
      if (f == 0) {
        sto(w.w_work, qf, L(A, 0));
        L(B, 0) = s_ph0[tx];
        continue;
      }
      double wv = L(A, f);
      double wd = wv;
      if (damp) {
        double hk = ph1c[r] / kG;
        double sn = wd_sin(((0.5 * kPi) * (hk - hbot)) / hdepth);
        double ramp = sn * sn;
        double dampwt = hk >= hbot ? dampmag * ramp : 0.0;
        double mass_f = g.c1f[f] * mut + g.c2f[f];
        // below the damping layer dampwt = 0 and the saved w drops out
        // exactly (wv - 0 * w_save) / 1, so it is only read inside the layer
        double wsave = hk >= hbot ? ldo(s.w, qf) : 0.0;
        wd = (wv - (dampwt * mass_f) * wsave) / (1.0 + dampwt);
      }
      double mfs = g.c1f[f] * muts + g.c2f[f];
      double phn = L(D, f) + (((((msfty * 0.5) * dts) * kG) * eps_p) * wd) / mfs;
      sto(w.ph_work, qf, phn);
      sto(w.w_work, qf, wd);
      L(B, f) = phn;
    }
  }
  __syncthreads();

  // ---- P8: calc_p_rho_step ----
  {
    double pm1r[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      // first substep of a stage: pm1 is the step-0 pressure (calc_p_rho_wrf
      // step 0 sets pm1 = p), read from p before this kernel overwrites it
      if (valid && k < nz) pm1r[r] = ldo(first ? w.p : w.pm1, c8 + (unsigned)k * plane);
    }
    const double mu_work = muts - mut;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = lev(r);
      if (!(valid && k < nz)) continue;
      const unsigned q = c8 + (unsigned)k * plane;
      double mass_h = g.c1h[k] * muts + g.c2h[k];
      double safe_mass = safe_pos(mass_h, 1.0e-12);
      double mu_term = g.c1h[k] * mu_work;
      double alt = altr[r];
      double al = -(alt * mu_term + g.rdnw[k] * (L(B, k + 1) - L(B, k))) / safe_mass;
      double theta_1 = thc[r] - kT0;
      double ttr = kT0 + theta_1;
      double safe_t = safe_pos(ttr, 1.0e-6);
      double p = c2ar[r] * (((alt * (L(E, k) - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
      double ptmp = p;
      p = p + kSmdiv * (p - pm1r[r]);
      sto(w.al, q, al);
      sto(w.p, q, p);
      sto(w.pm1, q, ptmp);
    }
  }
#undef L
}

template <int TX, int KT, int R>
__global__ void __launch_bounds__(TX * KT, (R <= 4 ? 3 : 1))
    k_column(DGrid g, DNamelist nl, DState s, DWork w, double dts, int first) {
  extern __shared__ double lds[];
  b_column<TX, KT, R>(threadIdx.x % TX, threadIdx.x / TX, tile_id(g), lds, g, nl, s, w, dts, first);
}

// ---------------------------------------------------------------------------
// small_step_finish_wrf (core/small_step_finish.py:22-101)
// ---------------------------------------------------------------------------
__device__ inline double safe_den(double ms) { return fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12); }

// One thread per face column (u faces first, then v faces) and a chunk of
// kFinChunk levels (the block's y index); the 2-D factors are loaded once per chunk.
constexpr int kFinChunk = 8;

__device__ inline void fin_face_column(const DGrid& g, unsigned chunk, unsigned f, unsigned fplane8, double mus,
                                       double mu, double msf, const double* __restrict__ xw, double* __restrict__ x) {
  const unsigned f8 = f * 8u;
  const int k0 = chunk * kFinChunk, k1 = k0 + kFinChunk < g.nz ? k0 + kFinChunk : g.nz;
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    const unsigned q = (unsigned)k * fplane8 + f8;
    const double ms = g.c1h[k] * mus + g.c2h[k];
    const double mc = g.c1h[k] * mu + g.c2h[k];
    sto(x, q, (msf * ldo(xw, q) + ldo(x, q) * mc) / safe_den(ms));
  }
}

__device__ __forceinline__ void b_fin_uv(const Blk& B, DGrid g, DState s, DWork w) {
  const unsigned ncu = (unsigned)(g.ny * (g.nx + 1)), ncv = (unsigned)((g.ny + 1) * g.nx);
  const unsigned t = B.x * B.bdx + B.tx;
  if (t < ncu) fin_face_column(g, B.y, t, ncu * 8u, w.muus[t], w.muu[t], g.msfuy[t], w.u_work, s.u);
  else if (t < ncu + ncv) {
    const unsigned f = t - ncu;
    fin_face_column(g, B.y, f, ncv * 8u, w.muvs[f], w.muv[f], g.msfvx[f], w.v_work, s.v);
  }
}

__global__ void __launch_bounds__(256) k_fin_uv(DGrid g, DState s, DWork w) {
  b_fin_uv(hw_blk(), g, s, w);
}

// ---------------------------------------------------------------------------
// Column-march form of the acoustic column work, split in two kernels; each
// value is computed with the same operations, in the same order, as k_column,
// by one thread per column with no LDS and no barriers.
//
// k_column_mu (advance_mu_t): sweep up for dvdxi and the dmdt sum, the mu /
// muts / muave updates, then a second sweep up for the omega recurrence at
// level f (stored, all levels, to w.ww) and the theta update and t_2ave at
// k = f-1.  Sweep 1 stores dvdxi per level in w.rhs, which k_column_w
// overwrites at every level before it reads it, and sweep 2 reads it back
// instead of recomputing it from four u and v faces.
//
// A block is kMuRows rows of one 64-lane wave each.  Waves overlap by two
// lanes: lanes 1..62 own columns and lanes 0 and 63 load the columns either
// side, so every east/west neighbour value (theta, the east u face) comes
// from the adjacent lane by a DPP wave shift.  North/south values of sweep 2
// (theta, the north v face) come from the waves above and below through LDS,
// exchanged one level behind the loads; a wave at the block edge, or whose
// neighbour row wraps periodically, loads them itself.
// ---------------------------------------------------------------------------
constexpr int kMuRows = 4;
constexpr int kMuOwned = 62;  // columns owned per wave

template <class TT>
__device__ __forceinline__ void b_column_mu(const Blk& B, double* lds, DGrid g, DNamelist nl,
                                            DState s, DWork w, double dts, int first) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned nbx = (unsigned)(nx + kMuOwned - 1) / kMuOwned;
  const unsigned t = tile_id(g, B);
  const int tx = B.tx, ty = B.ty;
  const int i = (int)(t % nbx) * kMuOwned + tx - 1;  // -1 and >= nx on the halo lanes
  const int j = (int)(t / nbx) * kMuRows + ty;
  const bool owned = tx >= 1 && tx <= kMuOwned && i >= 0 && i < nx && j < ny;
  // every lane loads in-bounds: its mass column (wrapped in i, row clamped),
  // its west u face (face index clamped to 0..nx) and its south v face
  const int ic = wrapi(i, nx), jr = j < ny ? j : ny - 1;
  const unsigned plane = (unsigned)((size_t)ny * nx) * 8u;
  const unsigned uplane = (unsigned)((size_t)ny * (nx + 1)) * 8u, vplane = (unsigned)((size_t)(ny + 1) * nx) * 8u;
  const unsigned c8 = (unsigned)i2(g, jr, ic) * 8u;
  const unsigned uw8 = (unsigned)i2u(g, jr, clampi(i, 0, nx)) * 8u;
  const unsigned vs8 = (unsigned)i2v(g, j <= ny ? j : ny, ic) * 8u;
  const unsigned vn8 = (unsigned)i2v(g, j + 1 <= ny ? j + 1 : ny, ic) * 8u;
  // wave-uniform: which north/south values this wave loads itself
  const bool load_vn = ty == kMuRows - 1;
  const bool load_thn = !(ty < kMuRows - 1 && j + 1 < ny);
  const bool load_ths = !(ty > 0 && j >= 1);
  const size_t c = c8 / 8u;
  const double msftx = g.msftx[c], msfty = g.msfty[c];
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double mut = w.mut[c];
  const size_t fvn = vn8 / 8u, fvs = vs8 / 8u, fuw = uw8 / 8u;
  const size_t fue = i2u(g, jr, clampi(i + 1, 0, nx));
  const double muv_n = w.muv[fvn], muv_s = w.muv[fvs];
  const double rmvx_n = 1.0 / g.msfvx[fvn], rmvx_s = 1.0 / g.msfvx[fvs];
  const double muu_e = w.muu[fue], muu_w = w.muu[fuw];
  const double msfuy_e = g.msfuy[fue], msfuy_w = g.msfuy[fuw];
  auto dvdxi_at = [&](int k, double vwn, double vws, double vn, double vs, double uwe, double uww, double ue,
                      double uw) {
    const double c1 = g.c1h[k], c2 = g.c2h[k];
    double v_n = vwn + ((c1 * muv_n + c2) * vn) * rmvx_n;
    double v_s = vws + ((c1 * muv_s + c2) * vs) * rmvx_s;
    double u_e = uwe + ((c1 * muu_e + c2) * ue) / msfuy_e;
    double u_w = uww + ((c1 * muu_w + c2) * uw) / msfuy_w;
    return (msftx * msfty) * (g.rdy * (v_n - v_s) + g.rdx * (u_e - u_w));
  };

  // ---- sweep 1: dmdt (advance_mu_t), then the mass updates ----
  double dmdt = 0.0;
#pragma unroll 4
  for (int k = 0; k < nz; ++k) {
    const unsigned kv = (unsigned)k * vplane, ku = (unsigned)k * uplane;
    const double uww = ldo(w.u_work, uw8 + ku), uw = ldo(s.u, uw8 + ku);
    // the east face of column i is the west face of column i + 1 (lane + 1)
    const double uwe = wave_shl1(0.0, uww), ue = wave_shl1(0.0, uw);
    const double dv = dvdxi_at(k, ldo(w.v_work, vn8 + kv), ldo(w.v_work, vs8 + kv), ldo(s.v, vn8 + kv),
                               ldo(s.v, vs8 + kv), uwe, uww, ue, uw);
    dmdt = (k == 0) ? g.dnw[0] * dv : dmdt + g.dnw[k] * dv;
    if (owned) sto(w.rhs, c8 + (unsigned)k * plane, dv);  // read back by sweep 2
  }
  const double p2_ww0 = w.ww[c];
  const double p2_wwtop = ldo(w.ww, c8 + (unsigned)nz * plane);
  double muts, muave;
  {
    const double mu_work_old = w.muts[c] - mut;
    const double mu_save = w.mu[c] - mu_work_old;
    const double mu_tendency = dmdt + 0.0;
    const double mu_work_new = mu_work_old + dts * mu_tendency;
    muts = mut + mu_work_new;
    muave = 0.5 * (eps_p * mu_work_new + eps_m * mu_work_old);
    if (owned) {
      w.mu[c] = mu_save + mu_work_new;
      s.mudf[c] = mu_tendency;
      w.muts[c] = muts;
      w.muave[c] = muave;
    }
  }

  const unsigned cjp = (unsigned)i2(g, wrapi(jr + 1, ny), ic) * 8u, cjm = (unsigned)i2(g, wrapi(jr - 1, ny), ic) * 8u;
  struct Mv {  // mass level data
    double vwn, vws, uwe, uww, dv;
    double thc, thn, ths, the, thw, thwk, tht, t2o;
  };
  double (*sh)[kMuRows][2][64] = reinterpret_cast<double (*)[kMuRows][2][64]>(lds);
  Mv m_1{};                           // level f-1
  double thc_2 = 0.0;                 // theta at f-2
  double b_1 = 0.0;                   // omega at f-1
  double wprev = p2_ww0, dv_1 = 0.0;  // omega prefix, dvdxi at f-1
  for (int f = 0; f <= nz; ++f) {
    Mv m{};
    if (f < nz) {
      const unsigned kp = (unsigned)f * plane, q = c8 + kp;
      const unsigned kv = (unsigned)f * vplane, ku = (unsigned)f * uplane;
      if (load_vn) m.vwn = ldo(w.v_work, vn8 + kv);
      m.vws = ldo(w.v_work, vs8 + kv);
      m.uww = ldo(w.u_work, uw8 + ku);
      m.dv = ldo(w.rhs, q);
      m.thc = ldo(s.theta, q);
      if (load_thn) m.thn = ldo(s.theta, cjp + kp);
      if (load_ths) m.ths = ldo(s.theta, cjm + kp);
      m.thwk = ldo(w.th_work, q);
      m.tht = ldt<TT>(w.tht, q);
      m.t2o = first ? 0.0 : ldo(w.t2ave, q);  // t_2ave starts each stage at 0
    }
    if (f >= 1) {
      // level f-1's neighbours: east/west from the adjacent lanes, north/south
      // from the adjacent waves (LDS, double-buffered by level)
      m_1.uwe = wave_shl1(0.0, m_1.uww);
      m_1.the = wave_shl1(0.0, m_1.thc);
      m_1.thw = wave_shr1(0.0, m_1.thc);
      double* mine = &sh[f & 1][ty][0][tx];
      mine[0] = m_1.thc;
      mine[64] = m_1.vws;
      __syncthreads();
      if (!load_thn) m_1.thn = sh[f & 1][ty + 1][0][tx];
      if (!load_vn) m_1.vwn = sh[f & 1][ty + 1][1][tx];
      if (!load_ths) m_1.ths = sh[f & 1][ty - 1][0][tx];
    }
    const unsigned qf = c8 + (unsigned)f * plane;
    const double rom = f < nz ? ldo(w.rom, qf) : 0.0;
    const double dv = f < nz ? m.dv : 0.0;
    // omega at f (levels 0 and nz carry to the next substep; the interior is
    // rebuilt each substep and read by k_column_w)
    double b0;
    if (f == 0) {
      b0 = wprev - rom;
      if (owned) sto(w.ww, qf, b0);
    } else if (f < nz) {
      const int k = f - 1;
      double incr = (g.dnw[f - 1] * ((g.c1h[k] * dmdt + dv_1) + g.c1h[k] * 0.0)) / msfty;
      double wk = wprev - incr;
      b0 = wk - rom;
      wprev = wk;
      if (owned) sto(w.ww, qf, b0);
    } else {
      b0 = p2_wwtop;
    }
    // theta update and t_2ave at k = f-1
    if (f >= 1 && owned) {
      const int k = f - 1;
      const double th = m_1.thc - kT0;
      double wd_hi = 0.0, wd_lo = 0.0;
      if (k + 1 < nz) {
        const int kk = k + 1;
        wd_hi = b0 * (g.fnm[kk] * (m.thc - kT0) + g.fnp[kk] * th);
      }
      if (k >= 1) wd_lo = b_1 * (g.fnm[k] * th + g.fnp[k] * (thc_2 - kT0));
      double v_flux = m_1.vwn * ((m_1.thn - kT0) + th) - m_1.vws * (th + (m_1.ths - kT0));
      double u_flux = m_1.uwe * ((m_1.the - kT0) + th) - m_1.uww * (th + (m_1.thw - kT0));
      double tendency = msftx * ((0.5 * g.rdy) * v_flux + (0.5 * g.rdx) * u_flux) + g.rdnw[k] * (wd_hi - wd_lo);
      const unsigned q = c8 + (unsigned)k * plane;
      double theta_i = m_1.thwk + ((msfty * dts) * m_1.tht);
      double th_new = theta_i - ((dts * msfty) * tendency);
      sto(w.th_work, q, th_new);
      double mass_h = g.c1h[k] * muts + g.c2h[k];
      double half = 0.5 * (eps_p * th_new + eps_m * m_1.t2o);
      double ttr = kT0 + th;
      double t2n = (half + (g.c1h[k] * muave) * kT0) / (mass_h * ttr);
      sto(w.t2ave, q, t2n);
    }
    b_1 = b0;
    thc_2 = m_1.thc;
    dv_1 = dv;
    m_1 = m;
  }
}

template <class TT>
__global__ void __launch_bounds__(64 * kMuRows) k_column_mu(DGrid g, DNamelist nl, DState s, DWork w, double dts,
                                                           int first) {
  __shared__ double sh[2][kMuRows][2][64];
  b_column_mu<TT>(hw_blk(), &sh[0][0][0][0], g, nl, s, w, dts, first);
}

// k_column_w (advance_w + calc_p_rho): sweep up for the phi-equation rhs at
// f-1 (stored to w.rhs) and the explicit w and Thomas forward sweep at f-2
// (stored to w.wfwd), reading omega and the new t_2ave from k_column_mu; then
// sweep down for the back substitution, damping, the phi finish and
// calc_p_rho (which needs the finished phi of levels k and k+1).
template <class TT>
__device__ __forceinline__ void b_column_w(const Blk& B, DGrid g, DNamelist nl, DState s,
                                           DWork w, double dts, int first) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t ncol = (size_t)ny * nx;
  const size_t c = (size_t)tile_id(g, B) * B.bdx + B.tx;
  if (c >= ncol) return;
  const int j = (int)(c / (size_t)nx), i = (int)(c % (size_t)nx);
  const unsigned plane = (unsigned)ncol * 8u;
  const unsigned c8 = (unsigned)c * 8u;
  const double msftx = g.msftx[c], msfty = g.msfty[c];
  const double msft_inv = 1.0 / msfty;
  const double eps_p = 1.0 + nl.epssm, eps_m = 1.0 - nl.epssm;
  const double mut = w.mut[c];
  const double muts = w.muts[c], muave = w.muave[c];
  // terrain surface w (advance_w.py:505-556, level 0 of the explicit w)
  double a0;
  {
    const int jn = clampi(j + 1, 0, ny - 1), js = clampi(j - 1, 0, ny - 1);
    const int ie = clampi(i + 1, 0, nx - 1), iw = clampi(i - 1, 0, nx - 1);
    auto PHB0 = [&](int jj, int ii) {
      const size_t q = i2(g, jj, ii);
      return s.ph_total[q] - s.ph_pert[q];
    };
    double ht = PHB0(j, i) / kG;
    double ht_n = PHB0(jn, i) / kG, ht_s = PHB0(js, i) / kG;
    double ht_e = PHB0(j, ie) / kG, ht_w = PHB0(j, iw) / kG;
    double dy_n = ht_n - ht, dy_s = ht - ht_s, dx_e = ht_e - ht, dx_w = ht - ht_w;
    double v_cf_n = (g.cf1 * s.v[iv(g, 0, j + 1, i)] + g.cf2 * s.v[iv(g, 1, j + 1, i)]) + g.cf3 * s.v[iv(g, 2, j + 1, i)];
    double v_cf_s = (g.cf1 * s.v[iv(g, 0, j, i)] + g.cf2 * s.v[iv(g, 1, j, i)]) + g.cf3 * s.v[iv(g, 2, j, i)];
    double u_cf_e = (g.cf1 * s.u[iu(g, 0, j, i + 1)] + g.cf2 * s.u[iu(g, 1, j, i + 1)]) + g.cf3 * s.u[iu(g, 2, j, i + 1)];
    double u_cf_w = (g.cf1 * s.u[iu(g, 0, j, i)] + g.cf2 * s.u[iu(g, 1, j, i)]) + g.cf3 * s.u[iu(g, 2, j, i)];
    a0 = ((msfty * 0.5) * g.rdy) * (dy_n * v_cf_n + dy_s * v_cf_s) +
         ((msftx * 0.5) * g.rdx) * (dx_e * u_cf_e + dx_w * u_cf_w);
  }

  const double c05geps = (0.5 * kG) * eps_m;
  const double c05dtg = (0.5 * dts) * kG;
  struct Wv {  // w level data (c2a, alt, t2n: mass level f; cqw: the moisture factor of face f)
    double ph1t, pht, wold, phw, rwt, ca, cal, ww, c2a, alt, t2n, cqw;
  };
  Wv w_1{}, w_2{}, w_3{};                   // levels f-1, f-2, f-3
  double d_1 = 0.0, d_2 = 0.0;              // phi rhs at f-2, f-3
  double ph1t_2 = 0.0;                      // ph1t at f-2
  double wf_1 = 0.0;                        // Thomas forward value at the previous explicit level
  // level f's loads are issued one iteration ahead, before level f-1's stores
  auto load_w = [&](int f, Wv& L) {
    if (f <= nz) {
      const unsigned q = c8 + (unsigned)f * plane;
      L.ph1t = ldo(w.ph1t, q);
      L.pht = ldt<TT>(w.ph_tend, q);
      L.wold = ldo(w.w_work, q);
      L.phw = ldo(w.ph_work, q);
      L.rwt = ldo(w.rw_tend, q);
      L.ca = ldo(w.ca, q);
      L.cal = ldo(w.calpha, q);
      L.ww = ldo(w.ww, q);
      // the post-pg_buoy_w cqw inside: cq1 = 1/(1 + raw cqw) with moisture
      L.cqw = w.cqwr ? 1.0 / (1.0 + ldo(w.cqwr, q)) : cqw_at(f, nz);
      if (f < nz) {
        L.c2a = ldo(w.c2a, q);
        L.alt = ldo(w.alt, q);
        L.t2n = ldo(w.t2ave, q);
      }
    }
  };
  // damp_opt 3: the damping layer starts at fdamp, the lowest face f >= 1
  // with ph1t(f) / g >= hbot, found on the way up; below it the damping
  // weight is 0 and w passes through unchanged, so the down sweep skips it
  const bool damp = (nl.damp_opt == 3 && nl.dampcoef > 0.0);
  const double dampmag = dts * nl.dampcoef, hdepth = nl.zdamp;
  const double htop = ldo(w.ph1t, c8 + (unsigned)nz * plane) / kG;
  const double hbot = htop - hdepth;
  int fdamp = nz + 1;
  Wv wv_next{};
  load_w(0, wv_next);
  for (int f = 0; f <= nz + 2; ++f) {
    const Wv wv = wv_next;
    wv_next = Wv{};
    load_w(f + 1, wv_next);
    if (damp && f >= 1 && f <= nz && fdamp > nz && wv.ph1t / kG >= hbot) fdamp = f;
    // phi rhs at fd = f-1 (advance_w.py:460-503)
    double d0 = 0.0;
    if (f >= 1 && f <= nz + 1) {
      const int fd = f - 1;
      if (fd > 0) {
        double rr = dts * (w_1.pht + c05geps * w_1.wold);
        if (fd < nz) {
          double wd_f1 = ((0.5 * (wv.ww + w_1.ww)) * g.rdnw[fd]) * (wv.ph1t - w_1.ph1t);
          double wd_f = ((0.5 * (w_1.ww + w_2.ww)) * g.rdnw[fd - 1]) * (w_1.ph1t - ph1t_2);
          double adv = dts * (g.fnm[fd] * wd_f1 + g.fnp[fd] * wd_f);
          rr = rr + (-adv);
        }
        double mass_f = g.c1f[fd] * mut + g.c2f[fd];
        rr = w_1.phw + (msfty * rr) / mass_f;
        if (fd == nz && nl.top_lid) rr = 0.0;
        d0 = rr;
      }
      sto(w.rhs, c8 + (unsigned)fd * plane, d0);
    }
    // explicit w (advance_w.py:505-556) and the Thomas forward sweep at fa = f-2
    if (f >= 2) {
      const int fa = f - 2;
      double a;
      if (fa == 0) {
        a = a0;
      } else if (fa == nz) {
        a = 0.0;  // rigid lid
      } else {
        double rw = w_2.rwt;
        if (nl.w_damping == 1) {
          double mass_f = g.c1f[fa] * mut + g.c2f[fa];
          double safe_mf = safe_pos(mass_f, 1.0e-12);
          double vert_cfl = fabs(((w_2.ww / safe_mf) * g.rdnw[fa]) * dts);
          double damp = (((-sgn(w_2.wold)) * kWAlpha) * (vert_cfl - kWBeta)) * mass_f;
          rw = rw + (vert_cfl > kWBeta ? damp : 0.0);
        }
        double wn = w_2.wold + dts * rw;
        double mh_u = safe_pos(g.c1h[fa] * mut + g.c2h[fa], 1.0e-12);
        double mh_l = safe_pos(g.c1h[fa - 1] * mut + g.c2h[fa - 1], 1.0e-12);
        double c2a_u = w_2.c2a, c2a_l = w_3.c2a;
        double coef_u = (c2a_u * g.rdnw[fa]) / mh_u;
        double coef_l = (c2a_l * g.rdnw[fa - 1]) / mh_l;
        double phk1 = w_1.phw, phk = w_2.phw, phkm = w_3.phw;
        double tA_u = coef_u * (eps_p * (d0 - d_1) + eps_m * (phk1 - phk));
        double tA_l = coef_l * (eps_p * (d_1 - d_2) + eps_m * (phk - phkm));
        double termA = ((msft_inv * w_2.cqw) * (c05dtg * g.rdn[fa])) * (tA_u - tA_l);
        double bu = (c2a_u * w_2.alt) * w_2.t2n;
        double bl = (c2a_l * w_3.alt) * w_3.t2n;
        double termB = ((dts * kG) * msft_inv) * (g.rdn[fa] * (bu - bl) - (g.c1f[fa] * muave));
        a = wn + (termA + termB);
      }
      double wf = fa == 0 ? a : (a - w_2.ca * wf_1) * w_2.cal;
      sto(w.wfwd, c8 + (unsigned)fa * plane, wf);
      wf_1 = wf;
    }
    d_2 = d_1;
    d_1 = d0;
    ph1t_2 = w_1.ph1t;
    w_3 = w_2;
    w_2 = w_1;
    w_1 = wv;
  }

  // ---- sweep 3: back substitution, damping, phi finish, calc_p_rho ----
  {
    const double mu_work = muts - mut;
    double w_next = 0.0, phn_next = 0.0;
    struct Dv {  // level f loads of the down sweep (mass-level fields for f < nz)
      double wf, cg, ph1t, rhs, phw0, alt, th, c2a, thw, pm1;
    };
    // level f-1's loads are issued before level f's stores
    auto load_d = [&](int f, Dv& L) {
      if (f < 0) return;
      const unsigned q = c8 + (unsigned)f * plane;
      L.wf = ldo(w.wfwd, q);
      if (f >= 1 && f < nz) L.cg = ldo(w.cgamma, q);
      if (f >= 1) {
        if (f >= fdamp) L.ph1t = ldo(w.ph1t, q);
        L.rhs = ldo(w.rhs, q);
      } else {
        L.phw0 = ldo(w.ph_work, q);  // the surface ph work is unchanged
      }
      if (f < nz) {
        L.alt = ldo(w.alt, q);
        L.th = ldo(s.theta, q);
        L.c2a = ldo(w.c2a, q);
        L.thw = ldo(w.th_work, q);
        // first substep of a stage: pm1 is the step-0 pressure (calc_p_rho_wrf
        // step 0 sets pm1 = p), read from p before it is overwritten here
        L.pm1 = ldo(first ? w.p : w.pm1, q);
      }
    };
    Dv dn_next{};
    load_d(nz, dn_next);
    for (int f = nz; f >= 0; --f) {
      const Dv dv = dn_next;
      dn_next = Dv{};
      load_d(f - 1, dn_next);
      const unsigned q = c8 + (unsigned)f * plane;
      double wv;
      if (f == nz || f == 0) wv = dv.wf;
      else wv = dv.wf - dv.cg * w_next;
      w_next = wv;
      double phn;
      if (f == 0) {
        sto(w.w_work, q, wv);
        phn = dv.phw0;
      } else {
        double wd = wv;
        if (f >= fdamp) {  // damp_opt 3 (fdamp stays past the lid without it)
          double hk = dv.ph1t / kG;
          double sn = wd_sin(((0.5 * kPi) * (hk - hbot)) / hdepth);
          double ramp = sn * sn;
          double dampwt = hk >= hbot ? dampmag * ramp : 0.0;
          double mass_f = g.c1f[f] * mut + g.c2f[f];
          // below the damping layer dampwt = 0 and the saved w drops out
          // exactly ((wv - 0 * w_save) / 1), so it is only read inside the layer
          double wsave = hk >= hbot ? ldo(s.w, q) : 0.0;
          wd = (wv - (dampwt * mass_f) * wsave) / (1.0 + dampwt);
        }
        double mfs = g.c1f[f] * muts + g.c2f[f];
        phn = dv.rhs + (((((msfty * 0.5) * dts) * kG) * eps_p) * wd) / mfs;
        sto(w.ph_work, q, phn);
        sto(w.w_work, q, wd);
      }
      if (f < nz) {  // calc_p_rho_step at mass level k = f
        const int k = f;
        double mass_h = g.c1h[k] * muts + g.c2h[k];
        double safe_mass = safe_pos(mass_h, 1.0e-12);
        double mu_term = g.c1h[k] * mu_work;
        double alt = dv.alt;
        double al = -(alt * mu_term + g.rdnw[k] * (phn_next - phn)) / safe_mass;
        double theta_1 = dv.th - kT0;
        double ttr = kT0 + theta_1;
        double safe_t = safe_pos(ttr, 1.0e-6);
        double p = dv.c2a * (((alt * (dv.thw - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
        double ptmp = p;
        p = p + kSmdiv * (p - dv.pm1);
        sto(w.al, q, al);
        sto(w.p, q, p);
        sto(w.pm1, q, ptmp);
      }
      phn_next = phn;
    }
  }
}

// 3 waves per SIMD: the kernel fits 168 VGPRs without spilling
template <class TT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_column_w(DGrid g, DNamelist nl,
                                                                                         DState s, DWork w,
                                                                                         double dts, int first) {
  b_column_w<TT>(hw_blk(), g, nl, s, w, dts, first);
}

// small_step_finish for w, ph, theta, p and mu (core/small_step_finish.py:22-101)
// followed by _refresh_grid_p_from_finished (operational_mode.py:2169-2231,
// diagnose_pressure_al_alt's base branch, acoustic_wrf.py:360-438).  One
// thread per column marches up the w levels: the finished ph of levels f-1
// and f stay in registers for mass level f-1, and the next level's loads are
// issued before the current level is computed.
__device__ __forceinline__ void b_fin_col(const Blk& B, DGrid g, DState s, DWork w) {
  const size_t ncol = (size_t)g.ny * g.nx;
  const size_t c = (size_t)tile_id(g, B) * B.bdx + B.tx;
  if (c >= ncol) return;
  const int nz = g.nz;
  const unsigned pl8 = (unsigned)ncol * 8u, c8 = (unsigned)c * 8u;
  const double msfty = g.msfty[c];
  const double muts = w.muts[c], mut = w.mut[c];
  const double mub = w.mub[c], mu_pert_new = w.mu[c];
  const bool hyps2 = g.hypsometric_opt == 2;
  {
    double mu_base = s.mu_total[c] - s.mu_pert[c];
    s.mu_pert[c] = mu_pert_new;
    s.mu_total[c] = mu_base + mu_pert_new;
  }
  struct Lv {
    double ww, wsave, phw, php, pht;           // w level f
    double th, thw, ptot, ppert, pwork, pb;    // mass level f
  };
  auto load = [&](int f, Lv& L) {
    const unsigned q = c8 + (unsigned)f * pl8;
    L.ww = ldo(w.w_work, q);
    L.wsave = ldo(s.w, q);
    L.phw = ldo(w.ph_work, q);
    L.php = ldo(s.ph_pert, q);
    L.pht = ldo(s.ph_total, q);
    if (f < nz) {
      L.th = ldo(s.theta, q);
      L.thw = ldo(w.th_work, q);
      L.ptot = ldo(s.p_total, q);
      L.ppert = ldo(s.p_pert, q);
      L.pwork = ldo(w.p, q);
      L.pb = ldo(w.pb, q);
    }
  };
  Lv cur{}, prev{};
  double pht_prev = 0.0, php_prev = 0.0;  // finished ph total / perturbation at f-1
  load(0, cur);
  for (int f = 0; f <= nz; ++f) {
    Lv nxt{};
    if (f < nz) load(f + 1, nxt);
    const unsigned q = c8 + (unsigned)f * pl8;
    // w and ph at level f
    double ms = g.c1f[f] * muts + g.c2f[f];
    double mc = g.c1f[f] * mut + g.c2f[f];
    sto(s.w, q, (msfty * cur.ww + cur.wsave * mc) / safe_den(ms));
    const double ph_pert = cur.phw + cur.php;
    const double ph_base = cur.pht - cur.php;
    const double ph_total = ph_base + ph_pert;
    sto(s.ph_total, q, ph_total);
    sto(s.ph_pert, q, ph_pert);
    if (f >= 1) {  // theta and p at mass level k = f - 1
      const int k = f - 1;
      const unsigned qk = q - pl8;
      double msh = g.c1h[k] * muts + g.c2h[k];
      double mch = g.c1h[k] * mut + g.c2h[k];
      double t_save = prev.th - kT0;
      double theta_pert = (prev.thw + t_save * mch) / safe_den(msh);
      double theta = theta_pert + kT0;
      sto(s.theta, qk, theta);
      double p_base = prev.ptot - prev.ppert;
      double p_total = p_base + prev.pwork;
      double p_pert = prev.pwork;
      double muts_r = mub + mu_pert_new;
      double phb_k = pht_prev - php_prev;
      double phb_k1 = ph_total - ph_pert;
      double mass_h_base = g.c1h[k] * mub + g.c2h[k];
      double denom_alb = g.dnw[k] * mass_h_base;
      double safe_da = safe_pos(denom_alb, 1.0e-12);
      double alb = -(phb_k1 - phb_k) / safe_da;
      double alt;
      if (hyps2) {
        const double dph_tot = (phb_k1 + ph_pert) - (phb_k + php_prev);
        const double alb_used = log_alpha(g, k, phb_k1 - phb_k, mub);
        alt = (log_alpha(g, k, dph_tot, muts_r) - alb_used) + alb_used;
      } else {
        double mass_w = g.c1h[k] * muts_r + g.c2h[k];
        double safe_mass = safe_pos(mass_w, 1.0e-12);
        double mu_term = (alb * g.c1h[k]) * mu_pert_new;
        double geo_term = g.rdnw[k] * (ph_pert - php_prev);
        double al = -(mu_term + geo_term) / safe_mass;
        alt = al + alb;
      }
      double ptot = p_from_theta_alt(theta, alt);
      double p_pert_new = ptot - prev.pb;
      double p_base2 = p_total - p_pert;
      sto(s.p_total, qk, p_base2 + p_pert_new);
      sto(s.p_pert, qk, p_pert_new);
    }
    pht_prev = ph_total;
    php_prev = ph_pert;
    prev = cur;
    cur = nxt;
  }
}

__global__ void __launch_bounds__(256) k_fin_col(DGrid g, DState s, DWork w) {
  b_fin_col(hw_blk(), g, s, w);
}

// Launch grids of the acoustic kernels, which the persistent ensemble step
// gives its virtual blocks too (64 x 4 tiles, or 256-thread blocks).
constexpr int kUvBx = 64, kUvBy = 4;  // k_uv's block: one wave per row of faces
inline unsigned uv_tiles(const DGrid& g) {
  return (unsigned)((g.nx + 1 + kUvBx - 2) / (kUvBx - 1)) * (unsigned)((g.ny + 1 + kUvBy - 1) / kUvBy);
}
inline unsigned column_mu_tiles(const DGrid& g) {
  return (unsigned)((g.nx + kMuOwned - 1) / kMuOwned) * (unsigned)((g.ny + kMuRows - 1) / kMuRows);
}
inline unsigned fin_chunks(const DGrid& g) { return (unsigned)((g.nz + kFinChunk - 1) / kFinChunk); }

template <int TX, int KT, int R>
hipError_t launch_column(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, double dts,
                         int first, hipStream_t st, Prof* prof) {
  const unsigned colgrid = (unsigned)((g.nx + TX - 1) / TX) * (unsigned)g.ny;
  size_t lds = (6 * (size_t)(g.nz + 1) * TX + 3 * TX) * sizeof(double);
  WD_TIMED(prof, KID_COLUMN, st, (k_column<TX, KT, R><<<colgrid, TX * KT, lds, st>>>(g, nl, s, w, dts, first)));
  return hipGetLastError();
}

// R = levels per thread, the smallest that covers the nz+1 w levels.
template <int TX, int KT>
hipError_t launch_column_levels(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                double dts, int first, hipStream_t st, Prof* prof) {
  switch ((g.nz + 1 + KT - 1) / KT) {
    case 1: return launch_column<TX, KT, 1>(g, nl, s, w, dts, first, st, prof);
    case 2: return launch_column<TX, KT, 2>(g, nl, s, w, dts, first, st, prof);
    case 3: return launch_column<TX, KT, 3>(g, nl, s, w, dts, first, st, prof);
    case 4: return launch_column<TX, KT, 4>(g, nl, s, w, dts, first, st, prof);
    case 5: return launch_column<TX, KT, 5>(g, nl, s, w, dts, first, st, prof);
    case 6: return launch_column<TX, KT, 6>(g, nl, s, w, dts, first, st, prof);
    case 7: return launch_column<TX, KT, 7>(g, nl, s, w, dts, first, st, prof);
    case 8: return launch_column<TX, KT, 8>(g, nl, s, w, dts, first, st, prof);
    default: return hipErrorInvalidValue;  // rejected by wd_create
  }
}

}  // namespace

// Column-tile shapes (TX columns x KT threads per column) compiled in.
const ColumnTile kColumnTiles[] = {{16, 16}, {8, 16}, {32, 8}, {16, 8}};
const int kNumColumnTiles = sizeof(kColumnTiles) / sizeof(kColumnTiles[0]);

namespace {

hipError_t launch_column_tile(int tile, const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              double dts, int first, hipStream_t st, Prof* prof) {
  if (tile == kColumnMarch) {
    const unsigned bs = 256;
    const size_t C = (size_t)g.ny * g.nx;
    const unsigned mu_tiles = column_mu_tiles(g);
    const dim3 mu_block(64, kMuRows);
    if (nl.mixed) {
      WD_TIMED(prof, KID_COLUMN_MU, st, (k_column_mu<float><<<mu_tiles, mu_block, 0, st>>>(g, nl, s, w, dts, first)));
      WD_TIMED(prof, KID_COLUMN_W, st, (k_column_w<float><<<blocks_for(C, bs), bs, 0, st>>>(g, nl, s, w, dts, first)));
    } else {
      WD_TIMED(prof, KID_COLUMN_MU, st, (k_column_mu<double><<<mu_tiles, mu_block, 0, st>>>(g, nl, s, w, dts, first)));
      WD_TIMED(prof, KID_COLUMN_W, st, (k_column_w<double><<<blocks_for(C, bs), bs, 0, st>>>(g, nl, s, w, dts, first)));
    }
    return hipGetLastError();
  }
  // the LDS-tiled kernel reads FP64 tendencies only
  if (nl.mixed) return hipErrorInvalidValue;
  switch (tile) {
    case 1: return launch_column_levels<8, 16>(g, nl, s, w, dts, first, st, prof);
    case 2: return launch_column_levels<32, 8>(g, nl, s, w, dts, first, st, prof);
    case 3: return launch_column_levels<16, 8>(g, nl, s, w, dts, first, st, prof);
    default: return launch_column_levels<16, 16>(g, nl, s, w, dts, first, st, prof);
  }
}

}  // namespace

hipError_t launch_acoustic_substep(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                   double dts, int first, int column_tile, hipStream_t st, Prof* prof) {
  constexpr int BX = kUvBx, BY = kUvBy;
  const unsigned uvgrid = uv_tiles(g);
  if (nl.mixed) {
    WD_TIMED(prof, KID_UV, st, (k_uv<BX, BY, float, false><<<uvgrid, dim3(BX, BY), 0, st>>>(g, s, w, dts)));
  } else if (w.cqu) {
    WD_TIMED(prof, KID_UV, st, (k_uv<BX, BY, double, true><<<uvgrid, dim3(BX, BY), 0, st>>>(g, s, w, dts)));
  } else {
    WD_TIMED(prof, KID_UV, st, (k_uv<BX, BY, double, false><<<uvgrid, dim3(BX, BY), 0, st>>>(g, s, w, dts)));
  }
  // the LDS-tiled kernel has no moisture factor: a moist state runs the column march
  return launch_column_tile(w.cqwr ? kColumnMarch : column_tile, g, nl, s, w, dts, first, st, prof);
}

hipError_t launch_substep_part(int part, const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                               double dts, int first, hipStream_t st) {
  if (nl.mixed) return hipErrorInvalidValue;
  constexpr int BX = kUvBx, BY = kUvBy;
  const unsigned bs = 256;
  const size_t C = (size_t)g.ny * g.nx;
  if (part == 0) {
    const unsigned uvgrid = uv_tiles(g);
    k_uv<BX, BY, double, false><<<uvgrid, dim3(BX, BY), 0, st>>>(g, s, w, dts);
  } else if (part == 1) {
    k_column_mu<double><<<column_mu_tiles(g), dim3(64, kMuRows), 0, st>>>(g, nl, s, w, dts, first);
  } else if (part == 2) {
    k_column_w<double><<<blocks_for(C, bs), bs, 0, st>>>(g, nl, s, w, dts, first);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_stage_finish(const DGrid& g, const DState& s, const DWork& w, hipStream_t st, Prof* prof) {
  const unsigned bs = 256;
  const size_t UVcols = (size_t)g.ny * (g.nx + 1) + (size_t)(g.ny + 1) * g.nx;
  const size_t C = (size_t)g.ny * g.nx;
  const unsigned kchunks = fin_chunks(g);
  WD_TIMED(prof, KID_FIN_UV, st, (k_fin_uv<<<dim3(blocks_for(UVcols, bs), kchunks), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_FIN_COL, st, (k_fin_col<<<blocks_for(C, bs), bs, 0, st>>>(g, s, w)));
  return hipGetLastError();
}

const char* kernel_name(int id) {
  static const char* names[KID_COUNT] = {
      "origin_copy", "k_ruv", "k_rom", "k_prep_col", "k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t",
      "k_prep_2d", "k_prep_faces", "k_prep_u", "k_prep_v", "k_col_setup",
      "k_uv", "k_column", "k_column_mu", "k_column_w", "k_fin_uv", "k_fin_col", "k_diff6", "k_smag", "k_limiter",
      "k_constk", "k_coriolis", "k_moist", "k_moist_update"};
  return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

}  // namespace wd
