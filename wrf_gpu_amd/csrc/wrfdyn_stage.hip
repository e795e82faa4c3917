// RK-stage setup kernels (once per RK stage): transport velocities, flux-form
// advection, large-step PGF, small_step_prep, calc_p_rho step 0, the stage
// rw_tend / ph_tend and the calc_coef_w tridiagonal coefficients.
//
// Every kernel computes each output value with the reference's expression
// order (see the file:line cited per kernel); build with -ffp-contract=off.
#include "wrfdyn_kernels.h"

namespace wd {

namespace {

// ---------------------------------------------------------------------------
// couple_velocities_periodic ru/rv, dynamics/flux_advection.py:394-428
// ---------------------------------------------------------------------------
// The element-wise kernels below give each thread one column (mass, u-face
// or v-face) and a chunk of kChunk levels (the block's y index): the 2-D factors are
// loaded once per chunk, the 3-D index is k * plane + column, and a wave's
// loads cover 64 consecutive elements of one level.
constexpr int kChunk = 8;

__device__ __forceinline__ void b_ruv(const Blk& B, DGrid g, const double* __restrict__ u,
                                      const double* __restrict__ v, const double* __restrict__ mu,
                                      double* __restrict__ ru, double* __restrict__ rv) {
  const unsigned ncol = (unsigned)(g.ny * g.nx);
  const unsigned c = B.x * B.bdx + B.tx;
  if (c >= ncol) return;
  const int j = (int)(c / (unsigned)g.nx), i = (int)(c - (unsigned)j * (unsigned)g.nx);
  const double muu = 0.5 * (mu[c] + mu[i2(g, j, wrapi(i - 1, g.nx))]);
  const double muv = 0.5 * (mu[c] + mu[i2(g, wrapi(j - 1, g.ny), i)]);
  const double msfuy = g.msfuy[i2u(g, j, i)], msfvx = g.msfvx[i2v(g, j, i)];
  const unsigned pl8 = ncol * 8u, upl8 = (unsigned)(g.ny * (g.nx + 1)) * 8u, vpl8 = (unsigned)((g.ny + 1) * g.nx) * 8u;
  const unsigned c8 = c * 8u, cu8 = (unsigned)i2u(g, j, i) * 8u;
  const int k0 = B.y * kChunk, k1 = k0 + kChunk < g.nz ? k0 + kChunk : g.nz;
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    const double mass_u = g.c1h[k] * muu + g.c2h[k];
    const double mass_v = g.c1h[k] * muv + g.c2h[k];
    sto(ru, (unsigned)k * pl8 + c8, mass_u * ldo(u, (unsigned)k * upl8 + cu8) / msfuy);
    sto(rv, (unsigned)k * pl8 + c8, mass_v * ldo(v, (unsigned)k * vpl8 + c8) / msfvx);
  }
}

__global__ void __launch_bounds__(256) k_ruv(DGrid g, const double* __restrict__ u, const double* __restrict__ v,
                                             const double* __restrict__ mu, double* __restrict__ ru,
                                             double* __restrict__ rv) {
  b_ruv(hw_blk(), g, u, v, mu, ru, rv);
}

// calc_ww_cp omega, dynamics/flux_advection.py:430-458 (one thread per column).
// Up to kRomRegLevels levels the column's divergence is read once and kept in
// registers for the second sweep; deeper grids read it twice.
constexpr int kRomRegLevels = 64;
__device__ __forceinline__ void b_rom(const Blk& B, DGrid g, const double* __restrict__ ru,
                                      const double* __restrict__ rv, double* __restrict__ rom) {
  const unsigned nbx = (g.nx + B.bdx - 1) / B.bdx;
  const unsigned t = tile_id(g, B);
  int i = (int)(t % nbx * B.bdx + B.tx);
  int j = (int)(t / nbx);
  if (i >= g.nx) return;
  int ip = wrapi(i + 1, g.nx), jp = wrapi(j + 1, g.ny);
  double msftx = g.msftx[i2(g, j, i)];
  auto div = [&](int k) {
    return (g.dnw[k] * msftx) * (g.rdx * (ru[im(g, k, j, ip)] - ru[im(g, k, j, i)]) +
                                 g.rdy * (rv[im(g, k, jp, i)] - rv[im(g, k, j, i)]));
  };
  double dmdt = 0.0, cum = 0.0;
  rom[im(g, 0, j, i)] = 0.0;
  if (g.nz <= kRomRegLevels) {
    // one pass over ru and rv: the column's divergence stays in registers
    // (every index is fixed once the loops are unrolled)
    double d[kRomRegLevels];
#pragma unroll
    for (int k = 0; k < kRomRegLevels; ++k)
      if (k < g.nz) {
        d[k] = div(k);
        dmdt += d[k];
      }
#pragma unroll
    for (int k = 0; k < kRomRegLevels; ++k)
      if (k < g.nz) {
        double incr = -((g.dnw[k] * g.c1h[k]) * dmdt) - d[k];
        cum = (k == 0) ? incr : cum + incr;
        if (k + 1 < g.nz) rom[im(g, k + 1, j, i)] = cum;
      }
  } else {
    for (int k = 0; k < g.nz; ++k) dmdt += div(k);
    for (int k = 0; k < g.nz; ++k) {
      double incr = -((g.dnw[k] * g.c1h[k]) * dmdt) - div(k);
      cum = (k == 0) ? incr : cum + incr;
      if (k + 1 < g.nz) rom[im(g, k + 1, j, i)] = cum;
    }
  }
  rom[im(g, g.nz, j, i)] = 0.0;
}

__global__ void __launch_bounds__(64) k_rom(DGrid g, const double* __restrict__ ru, const double* __restrict__ rv,
                                            double* __restrict__ rom) {
  b_rom(hw_blk(), g, ru, rv, rom);
}

// _vertical_flux_div_3 face flux (flux_advection.py:1362-1407) from a column
// window: qm2..qp1 = q at kf-2 .. kf+1, r = the transport at face kf.
__device__ inline double vface3(const DGrid& g, int kf, double qm2, double qm1, double q0, double qp1, double r) {
  const int nz = g.nz;
  if (kf == 0 || kf == nz) return 0.0;
  if (kf == 1) return r * (g.fnm[1] * q0 + g.fnp[1] * qm1);
  if (kf == nz - 1) return r * (g.fnm[nz - 1] * q0 + g.fnp[nz - 1] * qm1);
  return r * flux3(qm2, qm1, q0, qp1, -r);
}

// Mixed form: the face interpolant of vface3 applied to deviations d = q - qc
// (the face flux is r times it); 0 at the boundary faces, where r is 0.
__device__ inline float vface3_dev(const DGrid& g, int kf, float dm2, float dm1, float d0, float dp1, double r) {
  const int nz = g.nz;
  if (kf == 0 || kf == nz) return 0.0f;
  if (kf == 1) return (float)g.fnm[1] * d0 + (float)g.fnp[1] * dm1;
  if (kf == nz - 1) return (float)g.fnm[nz - 1] * d0 + (float)g.fnp[nz - 1] * dm1;
  return flux3f(dm2, dm1, d0, dp1, sgnf(-r));
}
// Mixed form: r_hi F_hi - r_lo F_lo for a cell of value qc, from the face
// interpolants of deviations (f_hi relative to qc, f_lo relative to the cell
// below, whose value is qb).
__device__ inline double vdiv_mixed(double qc, double qb, double r_lo, double r_hi, float f_lo_below, float f_hi) {
  const float f_lo = f_lo_below + (float)(qb - qc);
  const float dev = (float)r_hi * f_hi - (float)r_lo * f_lo;
  return (double)dev + qc * (r_hi - r_lo);
}

// dpn face value (large-step PGF, rk_addtend_dry.py:314-327, top_lid with
// cfn/cfn1) from the pair sums ps(k) = p(k, a) + p(k, b) of levels kk-2..kk
// (ps0 at kk, psm1 at kk-1, psm2 at kk-2); the bottom face takes levels 0..2.
__device__ inline double dpn_ls_top(const DGrid& g, double psm1, double psm2) {
  return 0.5 * (g.cfn * psm1 + g.cfn1 * psm2);
}
__device__ inline double dpn_ls_mid(const DGrid& g, int kk, double ps0, double psm1) {
  return 0.5 * (g.fnm[kk] * ps0 + g.fnp[kk] * psm1);
}
__device__ inline double dpn_ls_bot(const DGrid& g, double ps0, double ps1, double ps2) {
  return 0.5 * ((g.cf1 * ps0 + g.cf2 * ps1) + g.cf3 * ps2);
}

// The four tendency kernels below march up one column per thread: the
// vertical stencil (flux3 over k-2..k+1, the PGF pair sums over k-1..k+1)
// runs from register windows and the horizontal stencil (flux5 over +-3
// columns and rows) reads level k, all of whose loads are issued together.
// k_tend_w also issues level k+1's loads before computing level k; for the
// other three the registers that takes cost more occupancy than it gains.
// XCD-contiguous tile order keeps neighbouring rows in one L2.

constexpr int kTendURows = 4;

// advect_scalar_flux (flux_advection.py:482-568) on mass levels of q - qoff:
// theta - 300 into tht (plus the explicit-diffusion tendencies d6 and sm), or a
// moisture species (qoff 0, no diffusion) into its qt.
// TT: the tendency storage type; float selects the mixed-precision form.
// Blocks of kTendURows rows of one 64-lane wave each, waves overlapping by one
// lane either side (lanes 1..62 own columns): the +-1 columns and ru at i+1
// come from the adjacent lanes (DPP), the rows j-3..j+3 inside the block and
// rv at j+1 from the adjacent waves (LDS); the +-2, +-3 columns and rows
// outside the block (or wrapping periodically) are loaded.
template <class TT>
__device__ __forceinline__ void b_tend_t(const Blk& B, double* lds, DGrid g, DWork w,
                                         const double* __restrict__ th, double qoff, double* dst, const double* d6,
                                         const double* sm) {
  constexpr bool MX = sizeof(TT) == sizeof(float);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned nbx = (unsigned)(nx + 61) / 62;
  const unsigned tile = tile_id(g, B);
  const int tx = B.tx, ty = B.ty;
  const int i = (int)(tile % nbx) * 62 + tx - 1;  // outside 0..nx-1 on the halo lanes
  const int j = (int)(tile / nbx) * kTendURows + ty;
  const bool owned = tx >= 1 && tx <= 62 && i < nx && j < ny;
  const int ic = wrapi(i, nx), jr = j < ny ? j : ny - 1;
  const unsigned pl8 = (unsigned)((size_t)ny * nx) * 8u;
  const unsigned c8 = (unsigned)i2(g, jr, ic) * 8u;
  unsigned ox[7], oy[7];  // byte offsets of columns i-3..i+3 (row j) and rows j-3..j+3 (column i)
  bool lds_y[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    ox[d] = (unsigned)i2(g, jr, wrapi(ic + d - 3, nx)) * 8u;
    oy[d] = (unsigned)i2(g, wrapi(jr + d - 3, ny), ic) * 8u;
    const int r = ty + d - 3, jj = j + d - 3;
    lds_y[d] = r >= 0 && r < kTendURows && jj >= 0 && jj < ny;
  }
  const double msftx = g.msftx[c8 / 8u];
  // the scalar and rv at level k
  double (*sh)[kTendURows][2][64] = reinterpret_cast<double (*)[kTendURows][2][64]>(lds);
  struct Lv {
    double x[7], y[7];  // theta at columns i-3..i+3 and rows j-3..j+3 (x[3], y[3]: centre)
    double ru0, ru1, rv0, rv1;
  };
  // centre-column window of theta (k-1..k+3, raw) and rom (k+1, k+2)
  double tm1 = 0.0, t0 = ldo(th, c8), tp1 = ldo(th, pl8 + c8), tp2 = ldo(th, 2 * pl8 + c8);
  double rom1 = ldo(w.rom, pl8 + c8);
  double ru_n = ldo(w.ru, c8), rv_n = ldo(w.rv, c8);  // own column, one level ahead
  double vf_lo = 0.0;  // vertical face flux at k (0 at the surface)
  double rom0 = 0.0;   // mixed form: rom at face k, and its face interpolant (deviation from level k-1)
  float fdev_lo = 0.0f;
  Lv cur{};
  // One mass level.  Level 0 is peeled off the loop, which folds the
  // surface-face branch out of it (measured faster).
  auto level = [&](int k) __attribute__((always_inline)) {
    const unsigned o = (unsigned)k * pl8;
    const double tp3 = k + 3 < nz ? ldo(th, (unsigned)(k + 3) * pl8 + c8) : 0.0;
    const double rom2 = k + 2 <= nz ? ldo(w.rom, (unsigned)(k + 2) * pl8 + c8) : 0.0;
    cur.x[0] = ldo(th, o + ox[0]);
    cur.x[1] = ldo(th, o + ox[1]);
    cur.x[5] = ldo(th, o + ox[5]);
    cur.x[6] = ldo(th, o + ox[6]);
#pragma unroll
    for (int d = 0; d < 7; ++d)
      if (d != 3 && !lds_y[d]) cur.y[d] = ldo(th, o + oy[d]);
    if (!lds_y[4]) cur.rv1 = ldo(w.rv, o + oy[4]);
    cur.ru0 = ru_n;
    cur.rv0 = rv_n;
    if (k + 1 < nz) {
      ru_n = ldo(w.ru, o + pl8 + c8);
      rv_n = ldo(w.rv, o + pl8 + c8);
    }
    cur.x[3] = t0;
    cur.y[3] = t0;
    cur.x[2] = dpp_move<kDppWaveShr1>(0.0, t0);
    cur.x[4] = dpp_move<kDppWaveShl1>(0.0, t0);
    cur.ru1 = dpp_move<kDppWaveShl1>(0.0, cur.ru0);
    {
      double* mine = &sh[k & 1][ty][0][tx];
      mine[0] = t0;
      mine[64] = cur.rv0;
      __syncthreads();
#pragma unroll
      for (int d = 0; d < 7; ++d)
        if (d != 3 && lds_y[d]) cur.y[d] = sh[k & 1][ty + d - 3][0][tx];
      if (lds_y[4]) cur.rv1 = sh[k & 1][ty + 1][1][tx];
    }
    const double qm1 = tm1 - qoff, q0 = t0 - qoff, qp1 = tp1 - qoff, qp2 = tp2 - qoff;
    auto QX = [&](int sft) { return cur.x[sft + 3] - qoff; };
    auto QY = [&](int d) { return cur.y[d + 3] - qoff; };
    double t;
    if constexpr (MX) {
      t = ((-msftx) * g.rdx) * hdiv5_mixed(cur.x, QX(0), cur.ru0, cur.ru1);
      t = t - (msftx * g.rdy) * hdiv5_mixed(cur.y, QY(0), cur.rv0, cur.rv1);
      const float fdev_hi = vface3_dev(g, k + 1, (float)(qm1 - q0), 0.0f, (float)(qp1 - q0), (float)(qp2 - q0), rom1);
      t = t - g.rdnw[k] * vdiv_mixed(q0, qm1, rom0, rom1, fdev_lo, fdev_hi);
      fdev_lo = fdev_hi;
      rom0 = rom1;
    } else {
      double vel0 = cur.ru0, vel1 = cur.ru1;
      double fq0 = vel0 * flux5(QX(-3), QX(-2), QX(-1), QX(0), QX(1), QX(2), vel0);
      double fq1 = vel1 * flux5(QX(-2), QX(-1), QX(0), QX(1), QX(2), QX(3), vel1);
      t = ((-msftx) * g.rdx) * (fq1 - fq0);
      double vy0 = cur.rv0, vy1 = cur.rv1;
      double gy0 = vy0 * flux5(QY(-3), QY(-2), QY(-1), QY(0), QY(1), QY(2), vy0);
      double gy1 = vy1 * flux5(QY(-2), QY(-1), QY(0), QY(1), QY(2), QY(3), vy1);
      t = t - (msftx * g.rdy) * (gy1 - gy0);
      const double vf_hi = vface3(g, k + 1, qm1, q0, qp1, qp2, rom1);
      t = t - g.rdnw[k] * (vf_hi - vf_lo);
      vf_lo = vf_hi;
    }
    if (owned) stt<TT>(dst, o + c8, add_diff(t, d6, sm, o + c8));
    tm1 = t0;
    t0 = tp1;
    tp1 = tp2;
    tp2 = tp3;
    rom1 = rom2;
  };
  level(0);
  for (int k = 1; k < nz; ++k) level(k);

}

template <class TT>
__global__ void __launch_bounds__(64 * kTendURows) k_tend_t(DGrid g, DWork w, const double* __restrict__ th,
                                                             double qoff, double* dst, const double* d6,
                                                             const double* sm) {
  __shared__ double sh[2][kTendURows][2][64];
  b_tend_t<TT>(hw_blk(), &sh[0][0][0][0], g, w, th, qoff, dst, d6, sm);
}

// advect_w_flux + _vertical_flux_div_w with a rigid lid
// (flux_advection.py:1277-1329, 1410-1482) on w levels (nz+1, ny, nx); the
// transport is _mass_to_full_levels of ru/rv (flux_advection.py:1332-1348).
template <class TT>
__device__ __forceinline__ void b_tend_w(const Blk& B, DGrid g, DState s, DWork w) {
  constexpr bool MX = sizeof(TT) == sizeof(float);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t ncol = (size_t)ny * nx;
  const size_t c = (size_t)tile_id(g, B) * B.bdx + B.tx;
  if (c >= ncol) return;
  const int j = (int)(c / (size_t)nx), i = (int)(c % (size_t)nx);
  const unsigned pl8 = (unsigned)ncol * 8u;
  unsigned ox[7], oy[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    ox[d] = (unsigned)i2(g, j, wrapi(i + d - 3, nx)) * 8u;
    oy[d] = (unsigned)i2(g, wrapi(j + d - 3, ny), i) * 8u;
  }
  const unsigned c8 = ox[3], cip8 = ox[4], cjp8 = oy[4];
  const double msf = g.msftx[c];
  const bool do_y = ny > 1;
  const double* W = s.w;
  struct Lv {
    double x[7], y[7];
    double ru0, ru1, rv0, rv1;  // mass level f (f < nz)
  };
  auto load = [&](int f, Lv& L) {
    const unsigned o = (unsigned)f * pl8;
#pragma unroll
    for (int d = 0; d < 7; ++d) L.x[d] = ldo(W, o + ox[d]);
    if (do_y) {
#pragma unroll
      for (int d = 0; d < 7; ++d)
        if (d != 3) L.y[d] = ldo(W, o + oy[d]);
    }
    L.y[3] = L.x[3];
    if (f < nz) {
      L.ru0 = ldo(w.ru, o + c8);
      L.ru1 = ldo(w.ru, o + cip8);
      L.rv0 = ldo(w.rv, o + c8);
      L.rv1 = ldo(w.rv, o + cjp8);
    }
  };
  // centre-column window of w (f-1..f+3) and rom (f, f+1, f+2)
  double wm1 = 0.0, w0 = ldo(W, c8), wp1 = ldo(W, pl8 + c8), wp2 = ldo(W, 2 * pl8 + c8);
  double rom0 = ldo(w.rom, c8), rom1 = ldo(w.rom, pl8 + c8);
  double vfl_lo = 0.0;  // vflux(f); vflux(0) = 0
  double vf_lo = 0.0;   // mixed form: transport at face f and its interpolant (deviation from w(f-1))
  float fdev_lo = 0.0f;
  double ru0m = 0.0, ru1m = 0.0, rv0m = 0.0, rv1m = 0.0;  // ru/rv at mass level f-1
  Lv cur{};
  load(0, cur);
  // One w level.  Levels 0 and 1 are peeled off the loop, which folds the
  // surface-face and first-face branches out of it (measured faster).
  auto level = [&](int f) __attribute__((always_inline)) {
    const double wp3 = f + 3 <= nz ? ldo(W, (unsigned)(f + 3) * pl8 + c8) : 0.0;
    const double rom2 = f + 2 <= nz ? ldo(w.rom, (unsigned)(f + 2) * pl8 + c8) : 0.0;
    Lv nxt{};
    if (f + 1 <= nz) load(f + 1, nxt);
    auto m2f = [&](double r_f, double r_fm1) {
      if (f == 0) return r_f;
      if (f == nz) return r_fm1;
      return g.fnm[f] * r_f + g.fnp[f] * r_fm1;
    };
    auto WX = [&](int sft) { return cur.x[sft + 3]; };
    auto WY = [&](int d) { return cur.y[d + 3]; };
    double vx0 = m2f(cur.ru0, ru0m), vx1 = m2f(cur.ru1, ru1m);
    double t;
    if constexpr (MX) {
      t = ((-msf) * g.rdx) * hdiv5_mixed(cur.x, WX(0), vx0, vx1);
    } else {
      double fq0 = vx0 * flux5(WX(-3), WX(-2), WX(-1), WX(0), WX(1), WX(2), vx0);
      double fq1 = vx1 * flux5(WX(-2), WX(-1), WX(0), WX(1), WX(2), WX(3), vx1);
      t = ((-msf) * g.rdx) * (fq1 - fq0);
    }
    if (do_y) {
      double vy0 = m2f(cur.rv0, rv0m), vy1 = m2f(cur.rv1, rv1m);
      if constexpr (MX) {
        t = t - (msf * g.rdy) * hdiv5_mixed(cur.y, WY(0), vy0, vy1);
      } else {
        double g0 = vy0 * flux5(WY(-3), WY(-2), WY(-1), WY(0), WY(1), WY(2), vy0);
        double g1 = vy1 * flux5(WY(-2), WY(-1), WY(0), WY(1), WY(2), WY(3), vy1);
        t = t - (msf * g.rdy) * (g1 - g0);
      }
    }
    // vflux(f+1) from w(f-1..f+2) and VF(f+1) = 0.5 (rom(f+1) + rom(f))
    double vfl_hi = 0.0;
    const int kk = f + 1;
    if constexpr (MX) {
      double vf_hi = 0.0;
      float fdev_hi = 0.0f;
      if (kk >= 1 && kk < nz) {
        vf_hi = 0.5 * (rom1 + rom0);
        const float dp1 = (float)(wp1 - w0);
        if (kk == 1) fdev_hi = 0.5f * dp1;
        else fdev_hi = flux3f((float)(wm1 - w0), 0.0f, dp1, (float)(wp2 - w0), sgnf(-vf_hi));
      }
      if (f >= 1 && f < nz) t = t + (-g.rdn[f]) * vdiv_mixed(w0, wm1, vf_lo, vf_hi, fdev_lo, fdev_hi);
      vf_lo = vf_hi;
      fdev_lo = fdev_hi;
    } else {
      if (kk >= 1 && kk < nz) {
        const double vf = 0.5 * (rom1 + rom0);
        if (kk == 1) vfl_hi = (vf * 0.5) * (wp1 + w0);
        else vfl_hi = vf * flux3(wm1, w0, wp1, wp2, -vf);
      }
      if (f >= 1 && f < nz) t = t + (-g.rdn[f]) * (vfl_hi - vfl_lo);
      else t = t + 0.0;
    }
    stt<TT>(w.wt, (unsigned)f * pl8 + c8, add_diff(t, w.d6w, w.smw, (unsigned)f * pl8 + c8));
    wm1 = w0;
    w0 = wp1;
    wp1 = wp2;
    wp2 = wp3;
    rom0 = rom1;
    rom1 = rom2;
    vfl_lo = vfl_hi;
    if (f < nz) {
      ru0m = cur.ru0;
      ru1m = cur.ru1;
      rv0m = cur.rv0;
      rv1m = cur.rv1;
    }
    cur = nxt;
  };
  level(0);
  level(1);
  for (int f = 2; f <= nz; ++f) level(f);

}

template <class TT>
__global__ void __launch_bounds__(256) k_tend_w(DGrid g, DState s, DWork w) {
  b_tend_w<TT>(hw_blk(), g, s, w);
}

// advect_u_flux (flux_advection.py:1131-1202) + large_step_horizontal_pgf x
// (rk_addtend_dry.py:329-343, 357-361) on u faces (nz, ny, nx+1).  Face nx
// repeats face 0's advection (periodic); the PGF uses edge-padded pairs.
// WET: the moisture factor cqu scales the PGF; dry launches fold it to 1.
template <class TT, bool WET>
__device__ __forceinline__ void b_tend_u(const Blk& B, double* lds, DGrid g, DState s, DWork w) {
  constexpr bool MX = sizeof(TT) == sizeof(float);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // Blocks of kTendURows rows of one 64-lane wave each; waves overlap by one
  // lane either side (lanes 1..62 own u faces).  Each lane loads its own face
  // (wrapped), its own mass column ia and the PGF's right column; the +-1
  // face, ru/rv/rom at ia-1 and ia+1 and the PGF's left column come from the
  // adjacent lanes (DPP), the rows j-3..j+3 inside the block and rv at j+1
  // from the adjacent waves (LDS); the +-2, +-3 faces and rows outside the
  // block (or wrapping periodically) are loaded.
  const unsigned nbx = (unsigned)(nx + 1 + 61) / 62;
  const unsigned tile = tile_id(g, B);
  const int tx = B.tx, ty = B.ty;
  const int i = (int)(tile % nbx) * 62 + tx - 1;  // -1 and > nx on the halo lanes
  const int j = (int)(tile / nbx) * kTendURows + ty;
  const bool owned = tx >= 1 && tx <= 62 && i <= nx && j < ny;
  const int jr = j < ny ? j : ny - 1;
  const int ia = wrapi(i, nx);  // the collapsed periodic face (face nx repeats face 0)
  const unsigned pl8 = (unsigned)((size_t)ny * nx) * 8u, upl8 = (unsigned)((size_t)ny * (nx + 1)) * 8u;
  const bool do_y = ny > 1;
  unsigned ux[7], uy[7];  // u faces ia-3..ia+3 on row j; rows j-3..j+3 at face ia
  bool lds_y[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    ux[d] = (unsigned)i2u(g, jr, wrapi(ia + d - 3, nx)) * 8u;
    uy[d] = (unsigned)i2u(g, wrapi(jr + d - 3, ny), ia) * 8u;
    const int r = ty + d - 3, jj = j + d - 3;
    lds_y[d] = r >= 0 && r < kTendURows && jj >= 0 && jj < ny;
  }
  const unsigned ca8 = (unsigned)i2(g, jr, ia) * 8u;
  const unsigned jpa8 = (unsigned)i2(g, wrapi(jr + 1, ny), ia) * 8u;
  const double msf = g.msfux[i2u(g, jr, ia)];
  const int io = clampi(i, 0, nx);
  const unsigned fo8 = (unsigned)i2u(g, jr, io) * 8u;  // the output face
  // PGF columns (edge padded): this lane's right column; the left one is the
  // right column of lane - 1
  const int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
  const size_t cl = i2(g, jr, il), cr = i2(g, jr, ir);
  const unsigned cl8 = (unsigned)cl * 8u, cr8 = (unsigned)cr * 8u;
  const double msf_u = g.msfux[i2u(g, jr, io)] / g.msfuy[i2u(g, jr, io)];
  const double mub_l = s.mu_total[cl] - s.mu_pert[cl];
  const double mub_r = s.mu_total[cr] - s.mu_pert[cr];
  const double mt_l = mub_l + s.mu_pert[cl];
  const double mt_r = mub_r + s.mu_pert[cr];
  const double muu = 0.5 * (mt_l + mt_r);
  const double mu_l = s.mu_pert[cl], mu_r = s.mu_pert[cr];
  const double* U = s.u;
  const double* PP = s.p_pert;
  // u at face ia and rv at column ia, level k
  double (*sh)[kTendURows][2][64] = reinterpret_cast<double (*)[kTendURows][2][64]>(lds);
  struct Lv {
    double x[7], y[7];                          // u at faces ia-3..ia+3 / rows j-3..j+3
    double ru_a, ru_m, ru_p, rv_a, rv_m, rv_pa, rv_pm;  // transport pieces
    double pl1, pr1, phl1, phr1;                // p' and ph' at level k+1 (l, r)
    double altl, altr, ptl, ptr, all, alr, phpl, phpr;  // level k diagnostics (l, r)
  };
  auto shr = [](double v) { return dpp_move<kDppWaveShr1>(0.0, v); };
  auto shl = [](double v) { return dpp_move<kDppWaveShl1>(0.0, v); };
  // ru, rv at the own column, one level ahead (rv goes to LDS before it is used)
  double ru_n = ldo(w.ru, ca8), rv_n = do_y ? ldo(w.rv, ca8) : 0.0;
  auto load = [&](int k, Lv& L) {
    const unsigned o = (unsigned)k * pl8, ou = (unsigned)k * upl8;
    L.x[0] = ldo(U, ou + ux[0]);
    L.x[1] = ldo(U, ou + ux[1]);
    L.x[5] = ldo(U, ou + ux[5]);
    L.x[6] = ldo(U, ou + ux[6]);
    if (do_y) {
#pragma unroll
      for (int d = 0; d < 7; ++d)
        if (d != 3 && !lds_y[d]) L.y[d] = ldo(U, ou + uy[d]);
      if (!lds_y[4]) L.rv_pa = ldo(w.rv, o + jpa8);
    }
    L.ru_a = ru_n;
    L.rv_a = rv_n;
    if (k + 1 < nz) {
      ru_n = ldo(w.ru, o + pl8 + ca8);
      if (do_y) rv_n = ldo(w.rv, o + pl8 + ca8);
      L.pr1 = ldo(PP, o + pl8 + cr8);
    }
    L.phr1 = ldo(s.ph_pert, o + pl8 + cr8);
    L.altr = ldo(w.alt_raw, o + cr8);
    L.ptr = ldo(s.p_total, o + cr8);
    L.alr = ldo(w.al_ls, o + cr8);
    L.phpr = ldo(w.php_ls, o + cr8);
  };
  // level k's values from the adjacent lanes and waves
  auto exchange = [&](int k, Lv& L, double uc) {
    L.x[3] = uc;
    L.y[3] = uc;
    L.x[2] = shr(uc);
    L.x[4] = shl(uc);
    L.ru_m = shr(L.ru_a);
    L.ru_p = shl(L.ru_a);
    L.pl1 = shr(L.pr1);
    L.phl1 = shr(L.phr1);
    L.altl = shr(L.altr);
    L.ptl = shr(L.ptr);
    L.all = shr(L.alr);
    L.phpl = shr(L.phpr);
    double* mine = &sh[k & 1][ty][0][tx];
    mine[0] = uc;
    mine[64] = L.rv_a;
    __syncthreads();
    if (do_y) {
#pragma unroll
      for (int d = 0; d < 7; ++d)
        if (d != 3 && lds_y[d]) L.y[d] = sh[k & 1][ty + d - 3][0][tx];
      if (lds_y[4]) L.rv_pa = sh[k & 1][ty + 1][1][tx];
      L.rv_m = shr(L.rv_a);
      L.rv_pm = shr(L.rv_pa);
    }
  };
  // centre-face window of u (k-1..k+3) and the face omega (k+1, k+2)
  auto romq = [&](int kk) {
    const double ra = ldo(w.rom, (unsigned)kk * pl8 + ca8);
    return 0.5 * (ra + shr(ra));
  };
  const unsigned ua8 = ux[3];
  double qm1 = 0.0, q0 = ldo(U, ua8), qp1 = ldo(U, upl8 + ua8), qp2 = ldo(U, 2 * upl8 + ua8);
  double r1 = romq(1);
  double vf_lo = 0.0;
  double r0 = 0.0;  // mixed form: face-k transport and interpolant (deviation from level k-1)
  float fdev_lo = 0.0f;
  // p' window at (l, r): level k (p0), k-1 (pm) ; ph' at level k
  double pl0 = ldo(PP, cl8), pr0 = ldo(PP, cr8), plm = 0.0, prm = 0.0;
  double phl0 = ldo(s.ph_pert, cl8), phr0 = ldo(s.ph_pert, cr8);
  double dpn_k;
  {
    const double ps0 = pl0 + pr0;
    const double ps1 = ldo(PP, pl8 + cl8) + ldo(PP, pl8 + cr8);
    const double ps2 = ldo(PP, 2 * pl8 + cl8) + ldo(PP, 2 * pl8 + cr8);
    dpn_k = dpn_ls_bot(g, ps0, ps1, ps2);
  }
  Lv cur{};
  for (int k = 0; k < nz; ++k) {
    const double qp3 = k + 3 < nz ? ldo(U, (unsigned)(k + 3) * upl8 + ua8) : 0.0;
    const double r2 = k + 2 <= nz ? romq(k + 2) : 0.0;
    load(k, cur);
    exchange(k, cur, q0);
    // --- advection on the collapsed periodic face ia ---
    auto UX = [&](int sft) { return cur.x[sft + 3]; };
    auto UY = [&](int d) { return cur.y[d + 3]; };
    double velx0 = 0.5 * (cur.ru_a + cur.ru_m);
    double velx1 = 0.5 * (cur.ru_p + cur.ru_a);
    double t;
    if constexpr (MX) {
      t = ((-msf) * g.rdx) * hdiv5_mixed(cur.x, UX(0), velx0, velx1);
      if (do_y) {
        double vy0 = 0.5 * (cur.rv_a + cur.rv_m);
        double vy1 = 0.5 * (cur.rv_pa + cur.rv_pm);
        t = t - (msf * g.rdy) * hdiv5_mixed(cur.y, UY(0), vy0, vy1);
      }
      const float fdev_hi = vface3_dev(g, k + 1, (float)(qm1 - q0), 0.0f, (float)(qp1 - q0), (float)(qp2 - q0), r1);
      t = t + (-g.rdnw[k]) * vdiv_mixed(q0, qm1, r0, r1, fdev_lo, fdev_hi);
      fdev_lo = fdev_hi;
      r0 = r1;
    } else {
      double fq0 = velx0 * flux5(UX(-3), UX(-2), UX(-1), UX(0), UX(1), UX(2), velx0);
      double fq1 = velx1 * flux5(UX(-2), UX(-1), UX(0), UX(1), UX(2), UX(3), velx1);
      t = ((-msf) * g.rdx) * (fq1 - fq0);
      if (do_y) {
        double vy0 = 0.5 * (cur.rv_a + cur.rv_m);
        double vy1 = 0.5 * (cur.rv_pa + cur.rv_pm);
        double g0 = vy0 * flux5(UY(-3), UY(-2), UY(-1), UY(0), UY(1), UY(2), vy0);
        double g1 = vy1 * flux5(UY(-2), UY(-1), UY(0), UY(1), UY(2), UY(3), vy1);
        t = t - (msf * g.rdy) * (g1 - g0);
      }
      const double vf_hi = vface3(g, k + 1, qm1, q0, qp1, qp2, r1);
      t = t + (-g.rdnw[k]) * (vf_hi - vf_lo);
      vf_lo = vf_hi;
    }
    // --- large-step horizontal PGF at face i ---
    const double pl1 = cur.pl1, pr1 = cur.pr1;
    double dpn_k1;
    if (k + 1 == nz) dpn_k1 = dpn_ls_top(g, pl0 + pr0, plm + prm);
    else dpn_k1 = dpn_ls_mid(g, k + 1, pl1 + pr1, pl0 + pr0);
    double mass_u = g.c1h[k] * muu + g.c2h[k];
    double ph_term = (cur.phr1 - cur.phl1) + (phr0 - phl0);
    double p_term = (cur.altl + cur.altr) * (pr0 - pl0);
    double pb_l = cur.ptl - pl0;
    double pb_r = cur.ptr - pr0;
    double pb_term = (cur.all + cur.alr) * (pb_r - pb_l);
    double dpx = (((msf_u * 0.5) * g.rdx) * mass_u) * ((ph_term + p_term) + pb_term);
    double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (mu_l + mu_r));
    dpx = dpx + ((msf_u * g.rdx) * (cur.phpr - cur.phpl)) * bracket;
    if (owned) {
      const unsigned off = (unsigned)k * upl8 + fo8;
      stt<TT>(w.ut, off, add_diff(t, w.d6u, w.smu, off) + (-(WET ? ldo(w.cqu, off) : 1.0)) * dpx);
    }
    qm1 = q0;
    q0 = qp1;
    qp1 = qp2;
    qp2 = qp3;
    r1 = r2;
    plm = pl0;
    prm = pr0;
    pl0 = pl1;
    pr0 = pr1;
    phl0 = cur.phl1;
    phr0 = cur.phr1;
    dpn_k = dpn_k1;
  }
}

template <class TT, bool WET>
__global__ void __launch_bounds__(64 * kTendURows) k_tend_u(DGrid g, DState s, DWork w) {
  __shared__ double sh[2][kTendURows][2][64];
  b_tend_u<TT, WET>(hw_blk(), &sh[0][0][0][0], g, s, w);
}

// advect_v_flux (flux_advection.py:1205-1274) + large_step_horizontal_pgf y
// (rk_addtend_dry.py:345-361) on v faces (nz, ny+1, nx).  Face ny repeats
// face 0's advection (periodic); the PGF uses edge-padded pairs.
// WET: the moisture factor cqv scales the PGF; dry launches fold it to 1.
// k_tend_v's LDS fields: v, ru, rv, and the six PGF north-row values
constexpr int kTendVFields = 9;

template <class TT, bool WET>
__device__ __forceinline__ void b_tend_v(const Blk& B, double* lds, DGrid g, DState s, DWork w) {
  constexpr bool MX = sizeof(TT) == sizeof(float);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // Blocks of kTendURows v-face rows of one 64-lane wave each; waves overlap
  // by one lane either side (lanes 1..62 own columns).  Each lane loads its
  // own face (collapsed row ja, column wrapped), its own mass column and the
  // PGF's north row; the +-1 columns come from the adjacent lanes (DPP), the
  // rows ja-3..ja+3 inside the block, ru/rv on row ja-1, rv on row ja+1 and
  // the PGF's south row from the adjacent waves (LDS); the +-2, +-3 columns
  // and rows outside the block (or wrapping differently) are loaded.
  const unsigned nbx = (unsigned)(nx + 61) / 62;
  const unsigned tile = tile_id(g, B);
  const int tx = B.tx, ty = B.ty;
  const int i = (int)(tile % nbx) * 62 + tx - 1;  // outside 0..nx-1 on the halo lanes
  const int j = (int)(tile / nbx) * kTendURows + ty;
  const bool owned = tx >= 1 && tx <= 62 && i < nx && j <= ny;
  const int jv = j <= ny ? j : ny;    // face row, clamped
  const int ja = (jv == ny) ? 0 : jv;  // the collapsed periodic face row (face ny repeats face 0)
  const int ic = wrapi(i, nx);
  const int jm = wrapi(ja - 1, ny), jp = wrapi(ja + 1, ny);
  const unsigned pl8 = (unsigned)((size_t)ny * nx) * 8u, upl8 = (unsigned)((size_t)(ny + 1) * nx) * 8u;
  unsigned ux[7], uy[7];  // v faces i-3..i+3 on row ja; rows ja-3..ja+3 at column i
  bool lds_y[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    ux[d] = (unsigned)i2v(g, ja, wrapi(ic + d - 3, nx)) * 8u;
    uy[d] = (unsigned)i2v(g, wrapi(ja + d - 3, ny), ic) * 8u;
    const int r = ty + d - 3, jj = j + d - 3;
    lds_y[d] = r >= 0 && r < kTendURows && jj >= 0 && jj < ny;
  }
  // wave-uniform: row ja-1 (ru, rv) from the wave below, row ja+1 (rv) from
  // the wave above, the PGF's south row from the wave below
  const bool lds_m = ty > 0 && j >= 1 && j <= ny;
  const bool lds_p = ty < kTendURows - 1 && j + 1 <= ny;
  const bool lds_l = ty > 0;
  const unsigned ca8 = (unsigned)i2(g, ja, ic) * 8u, cjm8 = (unsigned)i2(g, jm, ic) * 8u;
  const unsigned cjp8 = (unsigned)i2(g, jp, ic) * 8u;
  const double msfvy = g.msfvy[i2v(g, ja, ic)];
  const double ratio = g.msfvy[i2v(g, ja, ic)] / g.msfvx[i2v(g, ja, ic)];
  const unsigned fo8 = (unsigned)i2v(g, jv, ic) * 8u;  // the output face
  // PGF rows (edge padded)
  const int js = clampi(jv - 1, 0, ny - 1), jn = clampi(jv, 0, ny - 1);
  const size_t cl = i2(g, js, ic), cr = i2(g, jn, ic);
  const unsigned cl8 = (unsigned)cl * 8u, cr8 = (unsigned)cr * 8u;
  const double msf_v = g.msfvy[i2v(g, jv, ic)] / g.msfvx[i2v(g, jv, ic)];
  const double mub_s = s.mu_total[cl] - s.mu_pert[cl];
  const double mub_n = s.mu_total[cr] - s.mu_pert[cr];
  const double mt_s = mub_s + s.mu_pert[cl];
  const double mt_n = mub_n + s.mu_pert[cr];
  const double muv = 0.5 * (mt_s + mt_n);
  const double mu_s = s.mu_pert[cl], mu_n = s.mu_pert[cr];
  const double* U = s.v;
  const double* PP = s.p_pert;
  double (*sh)[kTendURows][kTendVFields][64] = reinterpret_cast<double (*)[kTendURows][kTendVFields][64]>(lds);
  struct Lv {
    double x[7], y[7];                          // v at columns i-3..i+3 / rows ja-3..ja+3
    double ru_a, ru_m, ru_p, ru_mp, rv_a, rv_m, rv_p;   // transport pieces
    double pl1, pr1, phl1, phr1;                // p' and ph' at level k+1 (l, r)
    double altl, altr, ptl, ptr, all, alr, phpl, phpr;  // level k diagnostics (l, r)
  };
  auto shl = [](double v) { return dpp_move<kDppWaveShl1>(0.0, v); };
  auto shr = [](double v) { return dpp_move<kDppWaveShr1>(0.0, v); };
  double ru_n = ldo(w.ru, ca8), rv_n = ldo(w.rv, ca8);  // own column, one level ahead
  auto load = [&](int k, Lv& L) {
    const unsigned o = (unsigned)k * pl8, ou = (unsigned)k * upl8;
    L.x[0] = ldo(U, ou + ux[0]);
    L.x[1] = ldo(U, ou + ux[1]);
    L.x[5] = ldo(U, ou + ux[5]);
    L.x[6] = ldo(U, ou + ux[6]);
#pragma unroll
    for (int d = 0; d < 7; ++d)
      if (d != 3 && !lds_y[d]) L.y[d] = ldo(U, ou + uy[d]);
    L.ru_a = ru_n;
    L.rv_a = rv_n;
    if (k + 1 < nz) {
      ru_n = ldo(w.ru, o + pl8 + ca8);
      rv_n = ldo(w.rv, o + pl8 + ca8);
    }
    if (!lds_m) {
      L.ru_m = ldo(w.ru, o + cjm8);
      L.rv_m = ldo(w.rv, o + cjm8);
    }
    if (!lds_p) L.rv_p = ldo(w.rv, o + cjp8);
    if (k + 1 < nz) {
      L.pr1 = ldo(PP, o + pl8 + cr8);
      if (!lds_l) L.pl1 = ldo(PP, o + pl8 + cl8);
    }
    L.phr1 = ldo(s.ph_pert, o + pl8 + cr8);
    L.altr = ldo(w.alt_raw, o + cr8);
    L.ptr = ldo(s.p_total, o + cr8);
    L.alr = ldo(w.al_ls, o + cr8);
    L.phpr = ldo(w.php_ls, o + cr8);
    if (!lds_l) {
      L.phl1 = ldo(s.ph_pert, o + pl8 + cl8);
      L.altl = ldo(w.alt_raw, o + cl8);
      L.ptl = ldo(s.p_total, o + cl8);
      L.all = ldo(w.al_ls, o + cl8);
      L.phpl = ldo(w.php_ls, o + cl8);
    }
  };
  // level k's values from the adjacent lanes and waves
  auto exchange = [&](int k, Lv& L, double vc) {
    L.x[3] = vc;
    L.y[3] = vc;
    L.x[2] = shr(vc);
    L.x[4] = shl(vc);
    double* mine = &sh[k & 1][ty][0][tx];
    mine[0 * 64] = vc;
    mine[1 * 64] = L.ru_a;
    mine[2 * 64] = L.rv_a;
    mine[3 * 64] = L.pr1;
    mine[4 * 64] = L.phr1;
    mine[5 * 64] = L.altr;
    mine[6 * 64] = L.ptr;
    mine[7 * 64] = L.alr;
    mine[8 * 64] = L.phpr;
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 7; ++d)
      if (d != 3 && lds_y[d]) L.y[d] = sh[k & 1][ty + d - 3][0][tx];
    if (lds_m) {
      L.ru_m = sh[k & 1][ty - 1][1][tx];
      L.rv_m = sh[k & 1][ty - 1][2][tx];
    }
    if (lds_p) L.rv_p = sh[k & 1][ty + 1][2][tx];
    if (lds_l) {
      const double* below = &sh[k & 1][ty - 1][0][tx];
      L.pl1 = below[3 * 64];
      L.phl1 = below[4 * 64];
      L.altl = below[5 * 64];
      L.ptl = below[6 * 64];
      L.all = below[7 * 64];
      L.phpl = below[8 * 64];
    }
    L.ru_p = shl(L.ru_a);
    L.ru_mp = shl(L.ru_m);
  };
  // centre-face window of v (k-1..k+3) and the face omega (k+1, k+2)
  auto romq = [&](int kk) {
    const unsigned o = (unsigned)kk * pl8;
    return 0.5 * (ldo(w.rom, o + ca8) + ldo(w.rom, o + cjm8));
  };
  const unsigned ua8 = ux[3];
  double qm1 = 0.0, q0 = ldo(U, ua8), qp1 = ldo(U, upl8 + ua8), qp2 = ldo(U, 2 * upl8 + ua8);
  double r1 = romq(1);
  double vf_lo = 0.0;
  double r0 = 0.0;  // mixed form: face-k transport and interpolant (deviation from level k-1)
  float fdev_lo = 0.0f;
  // p' window at (l, r): level k (p0), k-1 (pm) ; ph' at level k
  double pl0 = ldo(PP, cl8), pr0 = ldo(PP, cr8), plm = 0.0, prm = 0.0;
  double phl0 = ldo(s.ph_pert, cl8), phr0 = ldo(s.ph_pert, cr8);
  double dpn_k;
  {
    const double ps0 = pl0 + pr0;
    const double ps1 = ldo(PP, pl8 + cl8) + ldo(PP, pl8 + cr8);
    const double ps2 = ldo(PP, 2 * pl8 + cl8) + ldo(PP, 2 * pl8 + cr8);
    dpn_k = dpn_ls_bot(g, ps0, ps1, ps2);
  }
  Lv cur{};
  for (int k = 0; k < nz; ++k) {
    const double qp3 = k + 3 < nz ? ldo(U, (unsigned)(k + 3) * upl8 + ua8) : 0.0;
    const double r2 = k + 2 <= nz ? romq(k + 2) : 0.0;
    load(k, cur);
    exchange(k, cur, q0);
    // --- advection on the collapsed periodic face ja ---
    auto VX = [&](int sft) { return cur.x[sft + 3]; };
    auto VY = [&](int d) { return cur.y[d + 3]; };
    double vx0 = 0.5 * (cur.ru_a + cur.ru_m);
    double vx1 = 0.5 * (cur.ru_p + cur.ru_mp);
    double vy0 = 0.5 * (cur.rv_a + cur.rv_m);
    double vy1 = 0.5 * (cur.rv_p + cur.rv_a);
    double t, vd;
    if constexpr (MX) {
      t = ((-msfvy) * g.rdx) * hdiv5_mixed(cur.x, VX(0), vx0, vx1);
      t = t - (msfvy * g.rdy) * hdiv5_mixed(cur.y, VY(0), vy0, vy1);
      const float fdev_hi = vface3_dev(g, k + 1, (float)(qm1 - q0), 0.0f, (float)(qp1 - q0), (float)(qp2 - q0), r1);
      vd = (-g.rdnw[k]) * vdiv_mixed(q0, qm1, r0, r1, fdev_lo, fdev_hi);
      fdev_lo = fdev_hi;
      r0 = r1;
    } else {
      double fq0 = vx0 * flux5(VX(-3), VX(-2), VX(-1), VX(0), VX(1), VX(2), vx0);
      double fq1 = vx1 * flux5(VX(-2), VX(-1), VX(0), VX(1), VX(2), VX(3), vx1);
      t = ((-msfvy) * g.rdx) * (fq1 - fq0);
      double g0 = vy0 * flux5(VY(-3), VY(-2), VY(-1), VY(0), VY(1), VY(2), vy0);
      double g1 = vy1 * flux5(VY(-2), VY(-1), VY(0), VY(1), VY(2), VY(3), vy1);
      t = t - (msfvy * g.rdy) * (g1 - g0);
      const double vf_hi = vface3(g, k + 1, qm1, q0, qp1, qp2, r1);
      vd = (-g.rdnw[k]) * (vf_hi - vf_lo);
      vf_lo = vf_hi;
    }
    t = t + ratio * vd;
    // --- large-step PGF at face j (pairs south = l, north = r) ---
    const double pl1 = cur.pl1, pr1 = cur.pr1;
    double dpn_k1;
    if (k + 1 == nz) dpn_k1 = dpn_ls_top(g, pl0 + pr0, plm + prm);
    else dpn_k1 = dpn_ls_mid(g, k + 1, pl1 + pr1, pl0 + pr0);
    double mass_v = g.c1h[k] * muv + g.c2h[k];
    double ph_term = (cur.phr1 - cur.phl1) + (phr0 - phl0);
    double p_term = (cur.altl + cur.altr) * (pr0 - pl0);
    double pb_s = cur.ptl - pl0;
    double pb_n = cur.ptr - pr0;
    double pb_term = (cur.all + cur.alr) * (pb_n - pb_s);
    double dpy = (((msf_v * 0.5) * g.rdy) * mass_v) * ((ph_term + p_term) + pb_term);
    double bracket = g.rdnw[k] * (dpn_k1 - dpn_k) - 0.5 * (g.c1h[k] * (mu_s + mu_n));
    dpy = dpy + ((msf_v * g.rdy) * (cur.phpr - cur.phpl)) * bracket;
    if (owned) {
      const unsigned off = (unsigned)k * upl8 + fo8;
      stt<TT>(w.vt, off, add_diff(t, w.d6v, w.smv, off) + (-(WET ? ldo(w.cqv, off) : 1.0)) * dpy);
    }
    qm1 = q0;
    q0 = qp1;
    qp1 = qp2;
    qp2 = qp3;
    r1 = r2;
    plm = pl0;
    prm = pr0;
    pl0 = pl1;
    pr0 = pr1;
    phl0 = cur.phl1;
    phr0 = cur.phr1;
    dpn_k = dpn_k1;
  }
}

template <class TT, bool WET>
__global__ void __launch_bounds__(64 * kTendURows) k_tend_v(DGrid g, DState s, DWork w) {
  __shared__ double sh[2][kTendURows][kTendVFields][64];
  b_tend_v<TT, WET>(hw_blk(), &sh[0][0][0][0], g, s, w);
}

// ---------------------------------------------------------------------------
// small_step_prep_wrf, core/small_step_prep.py:216-394
// ---------------------------------------------------------------------------
__device__ __forceinline__ void b_prep_2d(const Blk& B, DGrid g, DState s, DWork w, int rk_step) {
  size_t n = (size_t)g.ny * g.nx;
  size_t c = (size_t)tile_id(g, B) * B.bdx + B.tx;
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

__global__ void k_prep_2d(DGrid g, DState s, DWork w, int rk_step) {
  b_prep_2d(hw_blk(), g, s, w, rk_step);
}

// _u_face_average_2d / _v_face_average_2d (small_step_prep.py:77-88)
__device__ __forceinline__ void b_prep_faces(const Blk& B, DGrid g, DWork w) {
  const int nx = g.nx, ny = g.ny;
  size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx;
  size_t idx = (size_t)tile_id(g, B) * B.bdx + B.tx;
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

__global__ void k_prep_faces(DGrid g, DWork w) {
  b_prep_faces(hw_blk(), g, w);
}

// coupled work velocity (small_step_prep.py:216-394) on one face column
// (u faces: f = j * (nx + 1) + i; v faces: f = j * nx + i)
__device__ inline void prep_face_column(const DGrid& g, unsigned chunk, unsigned f, unsigned fplane8, double mus,
                                        double mu, double msf, const double* __restrict__ x0,
                                        const double* __restrict__ x, double* __restrict__ xw) {
  const unsigned f8 = f * 8u;
  const int k0 = chunk * kChunk, k1 = k0 + kChunk < g.nz ? k0 + kChunk : g.nz;
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    const unsigned q = (unsigned)k * fplane8 + f8;
    const double mref = g.c1h[k] * mus + g.c2h[k];
    const double mcur = g.c1h[k] * mu + g.c2h[k];
    sto(xw, q, (mref * ldo(x0, q) - mcur * ldo(x, q)) / msf);
  }
}

__device__ __forceinline__ void b_prep_u(const Blk& B, DGrid g, DState s, DWork w) {
  const unsigned ncol = (unsigned)(g.ny * (g.nx + 1));
  const unsigned f = B.x * B.bdx + B.tx;
  if (f >= ncol) return;
  prep_face_column(g, B.y, f, ncol * 8u, w.muus[f], w.muu[f], g.msfuy[f], w.u0, s.u, w.u_work);
}

__global__ void __launch_bounds__(256) k_prep_u(DGrid g, DState s, DWork w) {
  b_prep_u(hw_blk(), g, s, w);
}

__device__ __forceinline__ void b_prep_v(const Blk& B, DGrid g, DState s, DWork w) {
  const unsigned ncol = (unsigned)((g.ny + 1) * g.nx);
  const unsigned f = B.x * B.bdx + B.tx;
  if (f >= ncol) return;
  prep_face_column(g, B.y, f, ncol * 8u, w.muvs[f], w.muv[f], g.msfvx[f], w.v0, s.v, w.v_work);
}

__global__ void __launch_bounds__(256) k_prep_v(DGrid g, DState s, DWork w) {
  b_prep_v(hw_blk(), g, s, w);
}

// ---------------------------------------------------------------------------
// Column prep, one thread per column marching up the w levels f = 0..nz; the
// mass level k = f - 1 is done once level f is loaded, so every state field is
// read once and the ph values of level k+1 come from registers:
//   w level f:   small_step_prep w_work, ph_work, ph_1 + phb (ph1t) and the
//                carried omega ends (core/small_step_prep.py:216-394)
//   mass level k: stage-entry EOS alt and _absolute_diagnostics al, php
//                (core/rk_addtend_dry.py:127-229, hypsometric_opt=1;
//                acoustic_wrf.py:354-358), small_step_prep th_work, pb, alt,
//                c2a, php and calc_p_rho_wrf step 0 al, p
//                (core/calc_p_rho.py:48-139)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void b_prep_col(const Blk& B, DGrid g, DState s, DWork w) {
  const size_t ncol = (size_t)g.ny * g.nx;
  const size_t c = (size_t)tile_id(g, B) * B.bdx + B.tx;
  if (c >= ncol) return;
  const int nz = g.nz;
  const unsigned pl8 = (unsigned)ncol * 8u, c8 = (unsigned)c * 8u;
  const double msfty = g.msfty[c];
  const double muts0 = w.muts0[c], mut = w.mut[c], mu_work0 = w.mu_work0[c];
  const double mu_pert = s.mu_pert[c], mu_total = s.mu_total[c];
  const double mub_d = mu_total - mu_pert;
  const double muts_d = mu_total + mu_pert;
  const bool hyps2 = g.hypsometric_opt == 2;
  struct Lv {
    double w0, wc, ph0, php, pht;   // w level f
    double th, ptot, ppert, th0;    // mass level f
  };
  auto load = [&](int f, Lv& L) {
    const unsigned q = c8 + (unsigned)f * pl8;
    L.w0 = ldo(w.w0, q);
    L.wc = ldo(s.w, q);
    L.ph0 = ldo(w.ph0, q);
    L.php = ldo(s.ph_pert, q);
    L.pht = ldo(s.ph_total, q);
    if (f < nz) {
      L.th = ldo(s.theta, q);
      L.ptot = ldo(s.p_total, q);
      L.ppert = ldo(s.p_pert, q);
      L.th0 = ldo(w.th0, q);
    }
  };
  Lv cur{}, prev{};
  load(0, cur);
  for (int f = 0; f <= nz; ++f) {
    Lv nxt{};
    if (f < nz) load(f + 1, nxt);
    const unsigned q = c8 + (unsigned)f * pl8;
    {  // w level f
      double mref = g.c1f[f] * muts0 + g.c2f[f];
      double mcur = g.c1f[f] * mut + g.c2f[f];
      sto(w.w_work, q, (mref * cur.w0 - mcur * cur.wc) / msfty);
      sto(w.ph_work, q, cur.ph0 - cur.php);
      const double phb = cur.pht - cur.php;
      sto(w.ph1t, q, cur.ph0 + phb);
      if (f == 0 || f == nz) sto(w.ww, q, ldo(w.rom, q));
    }
    if (f >= 1) {  // mass level k = f - 1 (level k in prev, level k+1 in cur)
      const int k = f - 1;
      const unsigned qk = q - pl8;
      // stage-entry EOS and _absolute_diagnostics
      const double alt_raw = alt_from_theta_p(prev.th, prev.ptot);
      sto(w.alt_raw, qk, alt_raw);
      {
        double phb_k = prev.pht - prev.php;
        double phb_k1 = cur.pht - cur.php;
        double mass_base = g.c1h[k] * mub_d + g.c2h[k];
        double safe_base = safe_pos(mass_base, 1.0e-12);
        double alb = ((-g.rdnw[k]) * (phb_k1 - phb_k)) / safe_base;
        if (hyps2) {
          // the state's mu_total is the column mass here (rk_addtend_dry.py:207)
          const double dph_tot = (phb_k1 + cur.php) - (phb_k + prev.php);
          sto(w.al_ls, qk, log_alpha(g, k, dph_tot, mu_total) - log_alpha(g, k, phb_k1 - phb_k, mub_d));
        } else {
          double mass_t = g.c1h[k] * muts_d + g.c2h[k];
          double safe_t = safe_pos(mass_t, 1.0e-12);
          double mu_term = g.c1h[k] * mu_pert;
          sto(w.al_ls, qk, -(alb * mu_term + g.rdnw[k] * (cur.php - prev.php)) / safe_t);
        }
        double pht_k = phb_k + prev.php;
        double pht_k1 = phb_k1 + cur.php;
        sto(w.php_ls, qk, 0.5 * (pht_k + pht_k1));
      }
      // small_step_prep (mass level) and calc_p_rho step 0
      double mref = g.c1h[k] * muts0 + g.c2h[k];
      double mcur = g.c1h[k] * mut + g.c2h[k];
      double th_ref = prev.th0 - kT0, th_cur = prev.th - kT0;
      double th_work = mref * th_ref - mcur * th_cur;
      sto(w.th_work, qk, th_work);
      double pb = prev.ptot - prev.ppert;
      sto(w.pb, qk, pb);
      double alt_ac = alt_raw > kMinAcousticAlt ? alt_raw : kMinAcousticAlt;
      sto(w.alt, qk, alt_ac);
      double c2a = (kCpovcv * (pb + prev.ppert)) / alt_ac;
      sto(w.c2a, qk, c2a);
      double phb_k = prev.pht - prev.php;
      double phb_k1 = cur.pht - cur.php;
      sto(w.php, qk, 0.5 * (((phb_k + phb_k1) + prev.php) + cur.php));
      double mass_h = g.c1h[k] * muts0 + g.c2h[k];
      double safe_mass = safe_pos(mass_h, 1.0e-12);
      double mu_term = g.c1h[k] * mu_work0;
      double phw_k = prev.ph0 - prev.php;
      double phw_k1 = cur.ph0 - cur.php;
      double al = -(alt_ac * mu_term + g.rdnw[k] * (phw_k1 - phw_k)) / safe_mass;
      double ttr = kT0 + th_ref;
      double safe_t = safe_pos(ttr, 1.0e-6);
      double p = c2a * (((alt_ac * (th_work - mu_term * th_ref)) / (safe_mass * safe_t)) - al);
      sto(w.al, qk, al);
      sto(w.p, qk, p);
      // pm1 = p and t_2ave = 0 (calc_p_rho_wrf step 0, small_step_prep) are
      // not stored: the stage's first substep reads p for pm1 and starts
      // t_2ave at 0.
    }
    prev = cur;
    cur = nxt;
  }
}

__global__ void __launch_bounds__(256) k_prep_col(DGrid g, DState s, DWork w) {
  b_prep_col(hw_blk(), g, s, w);
}

// ---------------------------------------------------------------------------
// Stage column setup, one thread per column marching up the faces:
//  rw_tend = pg_buoy_w_moist(qtot=0) + advect_w + curvature
//    (core/advance_w.py:209-290; operational_mode.py:1730-1822)
//  ph_tend = rhs_ph_wrf order 2 periodic, rigid lid (core/rhs_ph.py:175-398)
//  a/alpha/gamma = calc_coef_w_wrf_coefficients (dynamics/acoustic_wrf.py:783-852)
// Level f-1's coupled face velocities and ph_1 totals come from registers;
// level f+1's loads are issued before level f is computed.
//
// A block is kColRows rows of one 64-lane wave each; waves overlap by two
// lanes (lanes 1..62 own columns, lanes 0 and 63 load the columns either
// side).  At level f each lane's ph_1 total, west u face and south v face are
// exchanged: east/west from the adjacent lanes (DPP), north/south from the
// adjacent waves (LDS); a wave at the block edge, or whose neighbour row
// wraps periodically, loads them itself.
// ---------------------------------------------------------------------------
constexpr int kColRows = 4;
constexpr int kColOwned = 62;

// WET: the state carries moisture (the raw cqw is read); a dry launch folds cq1 = 1, cq2 = 0
template <class TT, bool WET>
__device__ __forceinline__ void b_col_setup(const Blk& B, double* lds, DGrid g, DNamelist nl,
                                            DState s, DWork w, double dts) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned nbx = (unsigned)(nx + kColOwned - 1) / kColOwned;
  const unsigned tile = tile_id(g, B);
  const int tx = B.tx, ty = B.ty;
  const int i = (int)(tile % nbx) * kColOwned + tx - 1;  // -1 and >= nx on the halo lanes
  const int j = (int)(tile / nbx) * kColRows + ty;
  const bool owned = tx >= 1 && tx <= kColOwned && i >= 0 && i < nx && j < ny;
  // every lane loads in-bounds: its mass column (wrapped in i, row clamped),
  // its west u face (face index clamped to 0..nx) and its south v face
  const int ic = wrapi(i, nx), jr = j < ny ? j : ny - 1;
  const size_t ncol = (size_t)ny * nx;
  const size_t c = i2(g, jr, ic);
  const unsigned pl8 = (unsigned)ncol * 8u;
  const unsigned upl8 = (unsigned)((size_t)ny * (nx + 1)) * 8u, vpl8 = (unsigned)((size_t)(ny + 1) * nx) * 8u;
  const unsigned c8 = (unsigned)c * 8u;
  const size_t fuw = i2u(g, jr, clampi(i, 0, nx)), fue = i2u(g, jr, clampi(i + 1, 0, nx));
  const size_t fvs = i2v(g, j <= ny ? j : ny, ic), fvn = i2v(g, j + 1 <= ny ? j + 1 : ny, ic);
  const unsigned uw8 = (unsigned)fuw * 8u, ue8 = (unsigned)fue * 8u;
  const unsigned vs8 = (unsigned)fvs * 8u, vn8 = (unsigned)fvn * 8u;
  const unsigned cjp8 = (unsigned)i2(g, wrapi(jr + 1, ny), ic) * 8u, cjm8 = (unsigned)i2(g, wrapi(jr - 1, ny), ic) * 8u;
  // wave-uniform: which north/south values this wave loads itself
  const bool load_vn = ty == kColRows - 1;
  const bool load_n = !(ty < kColRows - 1 && j + 1 < ny);
  const bool load_s = !(ty > 0 && j >= 1);
  const double msfty = g.msfty[c];
  const double msft_inv = 1.0 / msfty;
  const double mut = w.mut[c], mub = w.mub[c];
  const double mu_p = mut - mub;
  const double mxy = g.msftx[c] / msfty;
  const double e_c = g.rot ? g.ecor[c] : 0.0, cosa_c = g.rot ? g.cosa[c] : 1.0, sina_c = g.rot ? g.sina[c] : 0.0;
  const double muu_w = w.muu[fuw], muu_e = w.muu[fue];
  const double msfuy_w = g.msfuy[fuw], msfuy_e = g.msfuy[fue];
  const double muv_s = w.muv[fvs], muv_n = w.muv[fvn];
  const double msfvx_s = g.msfvx[fvs], msfvx_n = g.msfvx[fvn];
  // moisture: the raw calc_cq cqw of face f (0 for a dry state) gives
  // cq1 = 1/(1 + cqw) and cq2 = cqw cq1 in rw_tend (pg_buoy_w,
  // core/advance_w.py:266-290) and cq1 in the implicit-w rows (calc_coef_w)
  double cof = (0.5 * dts * kG * (1.0 + nl.epssm));
  cof = cof * cof;
  const double lid_flag = nl.top_lid ? 0.0 : 1.0;
  auto MH = [&](int k) { return g.c1h[k] * mut + g.c2h[k]; };
  auto MF = [&](int k) { return g.c1f[k] * mut + g.c2f[k]; };
  auto PHT = [&](double pert, double total) { return pert + (total - pert); };
  // curvature pieces at mass level kk: ru, u, rv, v averaged to the column
  struct Cv {
    double ru_m, u_m, rv_m, v_m;
  };
  auto curv_at = [&](int kk, double uw, double ue, double vs, double vn) {
    Cv r;
    double mu_l = g.c1h[kk] * muu_w + g.c2h[kk];
    double mu_r = g.c1h[kk] * muu_e + g.c2h[kk];
    double ru_l = (mu_l * uw) / msfuy_w;
    double ru_r = (mu_r * ue) / msfuy_e;
    r.ru_m = 0.5 * (ru_l + ru_r);
    r.u_m = 0.5 * (uw + ue);
    double mv_s = g.c1h[kk] * muv_s + g.c2h[kk];
    double mv_n = g.c1h[kk] * muv_n + g.c2h[kk];
    double rv_s = (mv_s * vs) / msfvx_s;
    double rv_n = (mv_n * vn) / msfvx_n;
    r.rv_m = 0.5 * (rv_s + rv_n);
    r.v_m = 0.5 * (vs + vn);
    return r;
  };
  struct Lv {
    double pp, wt, uw, ue, vs, vn, romp, phpp, phtp, ww, c2a;
    double pnp, pnt, psp, pst;  // ph' and ph total at f on rows j+1, j-1 (edge waves only)
    double cqr;                 // raw moisture cqw at face f
  };
  auto load = [&](int f, Lv& L) {
    const unsigned o = (unsigned)f * pl8, ou = (unsigned)f * upl8, ov = (unsigned)f * vpl8;
    L.pp = ldo(s.p_pert, o + c8);
    L.wt = ldt<TT>(w.wt, o + c8);
    L.uw = ldo(s.u, ou + uw8);
    L.vs = ldo(s.v, ov + vs8);
    if (load_vn) L.vn = ldo(s.v, ov + vn8);
    L.romp = ldo(w.rom, o + pl8 + c8);
    L.phpp = ldo(s.ph_pert, o + pl8 + c8);
    L.phtp = ldo(s.ph_total, o + pl8 + c8);
    L.ww = ldo(s.w, o + c8);
    L.c2a = ldo(w.c2a, o + c8);
    if constexpr (WET) L.cqr = ldo(w.cqwr, o + c8);
    if (load_n) {
      L.pnp = ldo(s.ph_pert, o + cjp8);
      L.pnt = ldo(s.ph_total, o + cjp8);
    }
    if (load_s) {
      L.psp = ldo(s.ph_pert, o + cjm8);
      L.pst = ldo(s.ph_total, o + cjm8);
    }
  };
  double (*sh)[kColRows][2][64] = reinterpret_cast<double (*)[kColRows][2][64]>(lds);
  // level 0
  double p_m = ldo(s.p_pert, c8);  // p' at f-1
  double uw_m = ldo(s.u, uw8), ue_m = ldo(s.u, ue8), vs_m = ldo(s.v, vs8), vn_m = ldo(s.v, vn8);
  Cv cv_m = curv_at(0, uw_m, ue_m, vs_m, vn_m);
  double rom_m = ldo(w.rom, c8), rom_0 = ldo(w.rom, pl8 + c8);                 // rom at f-1, f
  double pht_m = PHT(ldo(s.ph_pert, c8), ldo(s.ph_total, c8));                // ph_1 total at f-1
  double pht_0 = PHT(ldo(s.ph_pert, pl8 + c8), ldo(s.ph_total, pl8 + c8));    // at f
  double c2a_m = ldo(w.c2a, c8);
  if (owned) {
    sto(w.rw_tend, c8, 0.0 + ldt<TT>(w.wt, c8));
    stt<TT>(w.ph_tend, c8, 0.0);
    sto(w.ca, c8, 0.0);
    sto(w.calpha, c8, 1.0);
    sto(w.cgamma, c8, 0.0);
  }
  double gamma_prev = 0.0;
  double cqr_m = 0.0;  // raw moisture cqw of the previous face
  Lv cur{};
  load(1, cur);
  for (int f = 1; f < nz; ++f) {
    Lv nxt{};
    if (f + 1 < nz) load(f + 1, nxt);
    const unsigned q = c8 + (unsigned)f * pl8;
    // level f neighbours: east/west from the adjacent lanes, north/south from
    // the adjacent waves (LDS, double-buffered by level)
    double pht_n, pht_s;
    const double pht_e = dpp_move<kDppWaveShl1>(0.0, pht_0), pht_w = dpp_move<kDppWaveShr1>(0.0, pht_0);
    cur.ue = dpp_move<kDppWaveShl1>(0.0, cur.uw);
    {
      double* mine = &sh[f & 1][ty][0][tx];
      mine[0] = pht_0;
      mine[64] = cur.vs;
      __syncthreads();
      pht_n = load_n ? PHT(cur.pnp, cur.pnt) : sh[f & 1][ty + 1][0][tx];
      pht_s = load_s ? PHT(cur.psp, cur.pst) : sh[f & 1][ty - 1][0][tx];
      if (!load_vn) cur.vn = sh[f & 1][ty + 1][1][tx];
    }
    // (dry: cq1 = 1/(1 + 0) = 1 and cq2 = 0 exactly)
    const double cq1 = WET ? 1.0 / (1.0 + cur.cqr) : 1.0, cq2 = WET ? cur.cqr * cq1 : 0.0;
    // ---- rw_tend ----
    {
      double rw = (msft_inv * kG) *
                  (((cq1 * g.rdn[f]) * (cur.pp - p_m) - g.c1f[f] * mu_p) - cq2 * (g.c1f[f] * mub + g.c2f[f]));
      rw = rw + cur.wt;
      const Cv cv0 = curv_at(f, cur.uw, cur.ue, cur.vs, cur.vn);
      double ru_f = g.fnm[f] * cv0.ru_m + g.fnp[f] * cv_m.ru_m;
      double u_f = g.fnm[f] * cv0.u_m + g.fnp[f] * cv_m.u_m;
      double rv_f = g.fnm[f] * cv0.rv_m + g.fnp[f] * cv_m.rv_m;
      double v_f = g.fnm[f] * cv0.v_m + g.fnp[f] * cv_m.v_m;
      double curv = kReradius * (ru_f * u_f + (mxy * rv_f) * v_f);
      // cosine-Coriolis (operational_mode.py:1818-1821)
      const double cor = g.rot ? e_c * (cosa_c * ru_f - (mxy * sina_c) * rv_f) : 0.0;
      if (owned) sto(w.rw_tend, q, rw + (cor + curv));
      cv_m = cv0;
    }
    // ---- ph_tend (rhs_ph) ----
    const double pht_p = PHT(cur.phpp, cur.phtp);  // ph_1 total at f+1
    {
      double wd_f = ((0.5 * (cur.romp + rom_0)) * g.rdnw[f]) * (pht_p - pht_0);
      double wd_fm = ((0.5 * (rom_0 + rom_m)) * g.rdnw[f - 1]) * (pht_0 - pht_m);
      double term3 = g.fnm[f] * wd_f + g.fnp[f] * wd_fm;
      double t = -term3;
      double mass_f = g.c1f[f] * mut + g.c2f[f];
      double gw = ((mass_f * kG) * cur.ww) * msft_inv;
      t = t + gw;
      double pti = pht_0;
      double dn_ = pht_n - pti, ds_ = pti - pht_s;
      double vn = cur.vn + vn_m;
      double vs = cur.vs + vs_m;
      double mvn = g.c1f[f] * muv_n + g.c2f[f];
      double mvs = g.c1f[f] * muv_s + g.c2f[f];
      double fy = (mvn * vn) * dn_ + (mvs * vs) * ds_;
      t = t + (-((0.25 * g.rdy) * msft_inv)) * fy;
      double de_ = pht_e - pti, dw_ = pti - pht_w;
      double ue = cur.ue + ue_m;
      double uw = cur.uw + uw_m;
      double mue = g.c1f[f] * muu_e + g.c2f[f];
      double muw = g.c1f[f] * muu_w + g.c2f[f];
      double fx = (mue * ue) * de_ + (muw * uw) * dw_;
      t = t + (-((0.25 * g.rdx) * msft_inv)) * fx;
      if (owned) stt<TT>(w.ph_tend, q, t);
    }
    // ---- calc_coef_w ----
    {
      const int k = f;
      double ak;
      if (k == 1) {
        ak = 0.0;
      } else {
        int kb = k - 1;
        double denom = MH(kb) * MF(kb);
        ak = (((((-cq1) * cof) * g.rdn[k]) * g.rdnw[k - 1]) * c2a_m) / denom;
      }
      double du = MH(k) * MF(k);
      double dl = MH(k - 1) * MF(k);
      double dc = MH(k) * MF(k + 1);
      double b = 1.0 + ((cq1 * cof) * g.rdn[k]) * ((g.rdnw[k] * cur.c2a) / du + (g.rdnw[k - 1] * c2a_m) / dl);
      double cc = (((((-cq1) * cof) * g.rdn[k]) * g.rdnw[k]) * cur.c2a) / dc;
      double al = 1.0 / (b - ak * gamma_prev);
      double gm = cc * al;
      if (owned) {
        sto(w.ca, q, ak);
        sto(w.calpha, q, al);
        sto(w.cgamma, q, gm);
      }
      gamma_prev = gm;
    }
    p_m = cur.pp;
    uw_m = cur.uw;
    ue_m = cur.ue;
    vs_m = cur.vs;
    vn_m = cur.vn;
    rom_m = rom_0;
    rom_0 = cur.romp;
    pht_m = pht_0;
    pht_0 = pht_p;
    c2a_m = cur.c2a;
    if constexpr (WET) cqr_m = cur.cqr;
    cur = nxt;
  }
  // top face nz (p_m, c2a_m, cqr_m now hold level nz-1): the raw cqw of face
  // nz - 1 (advance_w.py:266-290)
  if (!owned) return;  // no barrier follows
  const double cq1 = WET ? 1.0 / (1.0 + cqr_m) : 1.0, cq2 = WET ? cqr_m * cq1 : 0.0;
  const unsigned qt = c8 + (unsigned)nz * pl8;
  double rwt = (msft_inv * kG) *
               (((((cq1 * 2.0) * g.rdnw[nz - 1]) * (-p_m)) - g.c1f[nz] * mu_p) - cq2 * (g.c1f[nz] * mub + g.c2f[nz]));
  sto(w.rw_tend, qt, rwt + ldt<TT>(w.wt, qt));
  stt<TT>(w.ph_tend, qt, 0.0);
  double top_denom_a = MH(nz - 1) * MF(nz - 1);
  double top_denom_b = MH(nz - 1) * MF(nz);
  double r2 = g.rdnw[nz - 1] * g.rdnw[nz - 1];
  double a_top = ((((-2.0 * cof) * r2) * c2a_m) * lid_flag) / top_denom_a;
  double b_top = 1.0 + (((2.0 * cof) * r2) * c2a_m) / top_denom_b;
  sto(w.ca, qt, a_top);
  sto(w.calpha, qt, 1.0 / (b_top - a_top * gamma_prev));
  sto(w.cgamma, qt, 0.0);
}

template <class TT, bool WET>
__global__ void __launch_bounds__(64 * kColRows) k_col_setup(DGrid g, DNamelist nl, DState s, DWork w, double dts) {
  __shared__ double sh[2][kColRows][2][64];
  b_col_setup<TT, WET>(hw_blk(), &sh[0][0][0][0], g, nl, s, w, dts);
}

}  // namespace

// ---------------------------------------------------------------------------
// 6th-order numerical filter (dynamics/explicit_diffusion.py:159-223 _dflux6,
// _sixth_axis, sixth_order_diffusion_tendency; applied at
// runtime/operational_mode.py:2870-2876): out = mass * sixth(f) for one of
// u (kind 0), v (1), w (2), theta (3), periodic along x and (with more than
// one row) y over the field's own extents, with the monotonic limiter; mass
// from the stage mu_total (face averages for u and v, :2768-2774).  The
// tendency kernels add it before the large-step PGF.
// ---------------------------------------------------------------------------
// The LDS-tiled stencil kernels: a 256-thread block (64 i x 4 j) covers a
// tile of 64 columns x kTileRows rows of one level, each thread kTileRows/4
// rows; the tile and its halo are staged in LDS with coalesced loads, each
// value once, and every stencil reads LDS.  One grid layer per level, so no
// thread divides to find its point.
constexpr int kFieldBx = 64, kFieldBy = 4, kTileRows = 16;
inline dim3 tile_grid(int ncol, int nrow, int nlev) {
  return dim3((unsigned)((ncol + kFieldBx - 1) / kFieldBx), (unsigned)((nrow + kTileRows - 1) / kTileRows),
              (unsigned)nlev);
}
// i wrapped into [0, n): the common in-range case without a division.
__device__ inline int wrapc(int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  const int r = i % n;
  return r < 0 ? r + n : r;
}

// coef (dflux_p1 - dflux_p0) of the seven values f[-3..3] of a line.
__device__ inline double d6_line(double fm3, double fm2, double fm1, double f0, double fp1, double fp2, double fp3,
                                 double coef) {
  double d0 = (10.0 * (f0 - fm1) - 5.0 * (fp1 - fm2)) + (fp2 - fm3);
  double d1 = (10.0 * (fp1 - f0) - 5.0 * (fp2 - fm1)) + (fp3 - fm2);
  const double g0 = f0 - fm1, g1 = fp1 - f0;
  if (d0 * g0 <= 0.0) d0 = 0.0;
  if (d1 * g1 <= 0.0) d1 = 0.0;
  return coef * (d1 - d0);
}

constexpr int kD6Halo = 3, kD6W = kFieldBx + 2 * kD6Halo, kD6H = kTileRows + 2 * kD6Halo;

// The mu of a point of a filtered or diffused field: u faces (kind 0) and v
// faces (1) take mu_total averaged to the face with the edge value at the
// first and last face, w levels (2) and mass points (3) mu itself.
__device__ inline double face_mu(const DGrid& g, int kind, const double* mu, int j, int i) {
  const int nx = g.nx, ny = g.ny;
  if (kind == 0)
    return i == 0 ? mu[(size_t)j * nx] : (i == nx ? mu[(size_t)j * nx + nx - 1]
                                                 : 0.5 * (mu[(size_t)j * nx + i - 1] + mu[(size_t)j * nx + i]));
  if (kind == 1)
    return j == 0 ? mu[i] : (j == ny ? mu[(size_t)(ny - 1) * nx + i]
                                     : 0.5 * (mu[(size_t)(j - 1) * nx + i] + mu[(size_t)j * nx + i]));
  return mu[(size_t)j * nx + i];
}

// Its dry mass on level k: c1f mu + c2f on w levels, c1h mu + c2h elsewhere.
__device__ inline double face_mass(const DGrid& g, int kind, const double* mu, int k, int j, int i) {
  const double m = face_mu(g, kind, mu, j, i);
  return kind == 2 ? g.c1f[k] * m + g.c2f[k] : g.c1h[k] * m + g.c2h[k];
}

__global__ void __launch_bounds__(256) k_diff6(DGrid g, const double* f, int kind, int nlev, int nrow, int ncol,
                                               const double* mu, double coef, double* out) {
  __shared__ double tile[kD6H][kD6W];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * kFieldBx, j0 = blockIdx.y * kTileRows, k = blockIdx.z;
  const double* lev = f + (size_t)k * nrow * ncol;
  for (int e = ty * kFieldBx + tx; e < kD6H * kD6W; e += kFieldBx * kFieldBy) {
    const int r = e / kD6W, c = e - r * kD6W;
    tile[r][c] = lev[(size_t)wrapc(j0 - kD6Halo + r, nrow) * ncol + wrapc(i0 - kD6Halo + c, ncol)];
  }
  __syncthreads();
  const int i = i0 + tx;
  if (i >= ncol) return;
  const int c = tx + kD6Halo;
  for (int r = ty; r < kTileRows; r += kFieldBy) {
    const int j = j0 + r;
    if (j >= nrow) break;
    const double mass = face_mass(g, kind, mu, k, j, i);
    const double* row = tile[r + kD6Halo];
    double t = d6_line(row[c - 3], row[c - 2], row[c - 1], row[c], row[c + 1], row[c + 2], row[c + 3], coef);
    if (nrow > 1) {
      const int q = r + kD6Halo;
      t = t + d6_line(tile[q - 3][c], tile[q - 2][c], tile[q - 1][c], tile[q][c], tile[q + 1][c], tile[q + 2][c],
                      tile[q + 3][c], coef);
    }
    out[((size_t)k * nrow + j) * ncol + i] = mass * t;
  }
}

void launch_diff6(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, hipStream_t st) {
  const dim3 bs(kFieldBx, kFieldBy);
  const double coef = nl.diff_6th_factor * 0.015625 / (2.0 * nl.dt);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  k_diff6<<<tile_grid(nx + 1, ny, nz), bs, 0, st>>>(g, s.u, 0, nz, ny, nx + 1, s.mu_total, coef, w.d6u);
  k_diff6<<<tile_grid(nx, ny + 1, nz), bs, 0, st>>>(g, s.v, 1, nz, ny + 1, nx, s.mu_total, coef, w.d6v);
  k_diff6<<<tile_grid(nx, ny, nz + 1), bs, 0, st>>>(g, s.w, 2, nz + 1, ny, nx, s.mu_total, coef, w.d6w);
  k_diff6<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.theta, 3, nz, ny, nx, s.mu_total, coef, w.d6t);
}

// ---------------------------------------------------------------------------
// Constant-K diffusion (const_nu_m2_s > 0; runtime/operational_mode.py:2878-2929):
// conservative_constant_k_diffusion_tendency (dynamics/explicit_diffusion.py:282-331)
// of u, v, w and theta into the smu..smt slots, which the tendency kernels
// add after the filter.  The divergence of mass K grad f, each face mass the
// mean of its two neighbours, periodic over each array's own extent (u over
// its nx + 1 faces, v over its ny + 1), no flux through the first and last
// level.  The layer thickness is one number for the grid, the mean of
// (ph_total[k+1] - ph_total[k]) / g and at least 1 m (:2756-2758), summed in a
// fixed order: k_ck_col sums each column upward, k_ck_dz (one block) folds
// the columns into kCkParts interleaved partial sums and adds those in order.
constexpr int kCkParts = 1024;

__global__ void __launch_bounds__(256) k_ck_col(int nz, size_t C, const double* ph, double* colsum) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double s = 0.0;
  for (int k = 0; k < nz; ++k) s = s + (ph[(size_t)(k + 1) * C + c] - ph[(size_t)k * C + c]) / kG;
  colsum[c] = s;
}

__global__ void __launch_bounds__(kCkParts) k_ck_dz(size_t C, double n, const double* colsum, double* dz) {
  __shared__ double part[kCkParts];
  const int t = threadIdx.x;
  double s = 0.0;
  for (size_t c = (size_t)t; c < C; c += kCkParts) s = s + colsum[c];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int q = 0; q < kCkParts; ++q) tot = tot + part[q];
    const double mean = tot / n;
    dz[0] = mean < 1.0 ? 1.0 : mean;  // NaN stays NaN, as jnp.maximum
  }
}

// LDS-tiled like k_diff6: the level's field and face mu over the tile and a
// one-point halo, staged once with coalesced loads; the levels above and
// below are read at the thread's own point.
constexpr int kCkW = kFieldBx + 2, kCkH = kTileRows + 2;

__global__ void __launch_bounds__(256) k_constk(DGrid g, const double* f, int kind, int nlev, int nrow, int ncol,
                                                const double* mu, double K, const double* dzp, double* out) {
  __shared__ double tf[kCkH][kCkW], tm[kCkH][kCkW];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * kFieldBx, j0 = blockIdx.y * kTileRows, k = blockIdx.z;
  const size_t P = (size_t)nrow * ncol;
  const double* lev = f + (size_t)k * P;
  for (int e = ty * kFieldBx + tx; e < kCkH * kCkW; e += kFieldBx * kFieldBy) {
    const int r = e / kCkW, c = e - r * kCkW;
    const int jj = wrapc(j0 - 1 + r, nrow), ii = wrapc(i0 - 1 + c, ncol);
    tf[r][c] = lev[(size_t)jj * ncol + ii];
    tm[r][c] = face_mu(g, kind, mu, jj, ii);
  }
  __syncthreads();
  const int i = i0 + tx;
  if (i >= ncol) return;
  const bool wl = kind == 2;
  const double* c1 = wl ? g.c1f : g.c1h;
  const double* c2 = wl ? g.c2f : g.c2h;
  const double c1k = c1[k], c2k = c2[k];
  const double dx = g.dx, dy = g.dy, dz = dzp[0];
  const int c = tx + 1;
  for (int r = ty; r < kTileRows; r += kFieldBy) {
    const int j = j0 + r;
    if (j >= nrow) break;
    const int q = r + 1;
    const double f0 = tf[q][c], mu0 = tm[q][c];
    const double m0 = c1k * mu0 + c2k;
    const double me = c1k * tm[q][c + 1] + c2k, mw = c1k * tm[q][c - 1] + c2k;
    const double fe = (0.5 * (m0 + me)) * K * (tf[q][c + 1] - f0) / dx;
    const double fw = (0.5 * (mw + m0)) * K * (f0 - tf[q][c - 1]) / dx;
    double t = (fe - fw) / dx;
    if (nrow > 1) {
      const double mn = c1k * tm[q + 1][c] + c2k, ms = c1k * tm[q - 1][c] + c2k;
      const double fn = (0.5 * (m0 + mn)) * K * (tf[q + 1][c] - f0) / dy;
      const double fs = (0.5 * (ms + m0)) * K * (f0 - tf[q - 1][c]) / dy;
      t = t + (fn - fs) / dy;
    }
    if (nlev > 2) {
      const size_t p = (size_t)j * ncol + i;
      double fu = 0.0, fd = 0.0;
      if (k + 1 < nlev) {
        const double mup = c1[k + 1] * mu0 + c2[k + 1];
        fu = (0.5 * (m0 + mup)) * K * (f[(size_t)(k + 1) * P + p] - f0) / dz;
      }
      if (k > 0) {
        const double mdn = c1[k - 1] * mu0 + c2[k - 1];
        fd = (0.5 * (mdn + m0)) * K * (f0 - f[(size_t)(k - 1) * P + p]) / dz;
      }
      t = t + (fu - fd) / dz;
    }
    out[(size_t)k * P + (size_t)j * ncol + i] = t;
  }
}

void launch_constk(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, hipStream_t st) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t C = (size_t)ny * nx;
  k_ck_col<<<blocks_for(C, 256), 256, 0, st>>>(nz, C, s.ph_total, w.ck_col);
  k_ck_dz<<<1, kCkParts, 0, st>>>(C, (double)((size_t)nz * C), w.ck_col, w.ck_dz);
  const dim3 bs(kFieldBx, kFieldBy);
  const double K = nl.const_nu;
  k_constk<<<tile_grid(nx + 1, ny, nz), bs, 0, st>>>(g, s.u, 0, nz, ny, nx + 1, s.mu_total, K, w.ck_dz, w.smu);
  k_constk<<<tile_grid(nx, ny + 1, nz), bs, 0, st>>>(g, s.v, 1, nz, ny + 1, nx, s.mu_total, K, w.ck_dz, w.smv);
  k_constk<<<tile_grid(nx, ny, nz + 1), bs, 0, st>>>(g, s.w, 2, nz + 1, ny, nx, s.mu_total, K, w.ck_dz, w.smw);
  k_constk<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.theta, 3, nz, ny, nx, s.mu_total, K, w.ck_dz, w.smt);
}

// ---------------------------------------------------------------------------
// 2-D Smagorinsky horizontal diffusion (diff_opt 1, km_opt 4;
// runtime/operational_mode.py:2946-2963).  k_smag_km: the eddy viscosity of
// smag2d_horizontal_km (dynamics/explicit_diffusion.py:1116-1217) from the
// horizontal deformations of horizontal_deformation_2d (:960-1113), which the
// reference calls with unit map factors and no slopes.  k_hdiff: the
// coordinate-surface flux divergence (_hdiff_coord_scalar, :1243-1284) of
// theta - 300 with K / prandtl and of u, v, w with K (:1287-1365), periodic
// over the mass extents; the u and v wrap faces take face 0's tendency.  The
// tendency kernels add it after the filter's, before the large-step PGF.
// ---------------------------------------------------------------------------
constexpr int kHdW = kFieldBx + 2, kHdH = kTileRows + 2;  // the tiles with a 1-point halo

// xkmh of the tile's cells from u and v staged with a 1-point halo (u, v
// periodic over the mass extents; tile (q, c) is cell (j0 + q - 1, i0 + c - 1)).
__global__ void __launch_bounds__(256) k_smag_km(DGrid g, const double* u, const double* v, double c_s, double* xkmh) {
  __shared__ double tu[kHdH][kHdW], tv[kHdH][kHdW];
  __shared__ double td12[kHdH][kHdW];  // D12 at the corners (q, c), q, c >= 1
  const int nx = g.nx, ny = g.ny;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * kFieldBx, j0 = blockIdx.y * kTileRows, k = blockIdx.z;
  for (int e = ty * kFieldBx + tx; e < kHdH * kHdW; e += kFieldBx * kFieldBy) {
    const int r = e / kHdW, c = e - r * kHdW;
    const int jj = wrapc(j0 - 1 + r, ny), ii = wrapc(i0 - 1 + c, nx);
    tu[r][c] = u[iu(g, k, jj, ii)];
    tv[r][c] = v[iv(g, k, jj, ii)];
  }
  __syncthreads();
  const double dx = g.dx, dy = g.dy;
  // D12 at the south-west corner of tile cell (q, c): du/dy + dv/dx, once
  // per corner the tile's cells use
  for (int e = ty * kFieldBx + tx; e < (kHdH - 1) * (kHdW - 1); e += kFieldBx * kFieldBy) {
    const int q = 1 + e / (kHdW - 1), cc = 1 + e % (kHdW - 1);
    const double dudy = (tu[q][cc] - tu[q - 1][cc]) / dy;
    const double dvdx = (tv[q][cc] - tv[q][cc - 1]) / dx;
    td12[q][cc] = dudy + dvdx;
  }
  __syncthreads();
  const int i = i0 + tx;
  if (i >= nx) return;
  const int c = tx + 1;
  for (int r = ty; r < kTileRows; r += kFieldBy) {
    const int j = j0 + r;
    if (j >= ny) break;
    const int q = r + 1;
    const double d11 = 2.0 * ((tu[q][c + 1] - tu[q][c]) / dx);
    const double d22 = 2.0 * ((tv[q + 1][c] - tv[q][c]) / dy);
    const double tmp = 0.25 * (((td12[q][c] + td12[q + 1][c]) + td12[q][c + 1]) + td12[q + 1][c + 1]);
    const double def2 = 0.25 * (d11 - d22) * (d11 - d22) + tmp * tmp;
    const double mlen = sqrt(dx * dy);
    const double km = c_s * c_s * mlen * mlen * sqrt(def2);
    const double cap = 10.0 * mlen;
    xkmh[((size_t)k * ny + j) * nx + i] = km > cap ? cap : km;
  }
}

// K of the w level f at column c: the mass-level K averaged to the face, the
// nearest level at faces 0 and nz (_xkmhd_to_w_levels, :1220-1240).
__device__ inline double smag_k_w(const double* xkmh, size_t plane, int nz, int f, size_t c) {
  if (f == 0) return xkmh[c];
  if (f == nz) return xkmh[(size_t)(nz - 1) * plane + c];
  return 0.5 * (xkmh[(size_t)f * plane + c] + xkmh[(size_t)(f - 1) * plane + c]);
}

// kind 0 u (nz, ny, nx+1), 1 v (nz, ny+1, nx), 2 w (nz+1, ny, nx),
// 3 theta (nz, ny, nx); out has the field's layout.  The field (theta - 300
// for theta), its K and its mass are staged per tile with a 1-point halo,
// periodic over the mass extents; the u and v wrap faces take face 0's value.
__global__ void __launch_bounds__(256) k_hdiff(DGrid g, const double* f, int kind, const double* xkmh,
                                               const double* mu, double* out) {
  __shared__ double tf[kHdH][kHdW], tk[kHdH][kHdW], tm[kHdH][kHdW];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int ncol = kind == 0 ? nx + 1 : nx, nrow = kind == 1 ? ny + 1 : ny;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * kFieldBx, j0 = blockIdx.y * kTileRows, k = blockIdx.z;
  const size_t plane = (size_t)ny * nx;
  const double* lev = f + (size_t)k * nrow * ncol;
  const double c1 = kind == 2 ? g.c1f[k] : g.c1h[k], c2 = kind == 2 ? g.c2f[k] : g.c2h[k];
  for (int e = ty * kFieldBx + tx; e < kHdH * kHdW; e += kFieldBx * kFieldBy) {
    const int r = e / kHdW, c = e - r * kHdW;
    const int jj = wrapc(j0 - 1 + r, ny), ii = wrapc(i0 - 1 + c, nx);
    const size_t cc = (size_t)jj * nx + ii;
    double fv, K, mf;
    if (kind == 0) {
      fv = lev[(size_t)jj * ncol + ii];
      K = xkmh[(size_t)k * plane + cc];
      mf = ii == 0 ? mu[(size_t)jj * nx] : 0.5 * (mu[cc - 1] + mu[cc]);
    } else if (kind == 1) {
      fv = lev[cc];
      K = xkmh[(size_t)k * plane + cc];
      mf = jj == 0 ? mu[ii] : 0.5 * (mu[cc - nx] + mu[cc]);
    } else if (kind == 2) {
      fv = lev[cc];
      K = smag_k_w(xkmh, plane, nz, k, cc);
      mf = mu[cc];
    } else {
      fv = lev[cc] - 300.0;
      K = xkmh[(size_t)k * plane + cc] / (1.0 / 3.0);
      mf = mu[cc];
    }
    tf[r][c] = fv;
    tk[r][c] = K;
    tm[r][c] = c1 * mf + c2;
  }
  __syncthreads();
  const int i = i0 + tx;
  if (i >= nx) return;
  const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
  const int c = tx + 1;
  for (int r = ty; r < kTileRows; r += kFieldBy) {
    const int j = j0 + r;
    if (j >= ny) break;
    const int q = r + 1;
    const double k_e = 0.5 * (tk[q][c + 1] + tk[q][c]), k_w = 0.5 * (tk[q][c] + tk[q][c - 1]);
    const double m_e = 0.5 * (tm[q][c + 1] + tm[q][c]), m_w = 0.5 * (tm[q][c] + tm[q][c - 1]);
    const double p = k_e * m_e * rdx, pw = k_w * m_w * rdx;
    const double f0 = tf[q][c];
    double t = rdx * (p * (tf[q][c + 1] - f0) - pw * (f0 - tf[q][c - 1]));
    if (ny > 1) {
      const double k_n = 0.5 * (tk[q + 1][c] + tk[q][c]), k_s = 0.5 * (tk[q][c] + tk[q - 1][c]);
      const double m_n = 0.5 * (tm[q + 1][c] + tm[q][c]), m_s = 0.5 * (tm[q][c] + tm[q - 1][c]);
      const double pn = k_n * m_n * rdy, ps = k_s * m_s * rdy;
      t = t + rdy * (pn * (tf[q + 1][c] - f0) - ps * (f0 - tf[q - 1][c]));
    }
    double* o = out + (size_t)k * nrow * ncol;
    o[(size_t)j * ncol + i] = t;
    if (kind == 0 && i == 0) o[(size_t)j * ncol + nx] = t;
    if (kind == 1 && j == 0) o[(size_t)ny * ncol + i] = t;
  }
}

void launch_smag(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, hipStream_t st) {
  const dim3 bs(kFieldBx, kFieldBy);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  k_smag_km<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.u, s.v, nl.c_s, w.xkmh);
  k_hdiff<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.u, 0, w.xkmh, s.mu_total, w.smu);
  k_hdiff<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.v, 1, w.xkmh, s.mu_total, w.smv);
  k_hdiff<<<tile_grid(nx, ny, nz + 1), bs, 0, st>>>(g, s.w, 2, w.xkmh, s.mu_total, w.smw);
  k_hdiff<<<tile_grid(nx, ny, nz), bs, 0, st>>>(g, s.theta, 3, w.xkmh, s.mu_total, w.smt);
}

// ---------------------------------------------------------------------------
// Limited scalar advection on the final RK stage (advect_scalar_flux_limited,
// dynamics/flux_advection.py:614-975): theta - 300 with scalar_adv_opt 1
// positive-definite or 2 monotonic (runtime/operational_mode.py:2819-2866),
// and each moisture species with moist_adv_opt (:3162-3230).  Donor-cell fluxes from the step
// origin's theta and mass, the antidiffusive remainder of the h5/v3 fluxes,
// the face fluxes (k_lim_faces, each face once), the per-cell scales
// (k_lim_scale), then the divergence of the limited fluxes (k_lim_tend),
// which replaces k_tend_t on that stage.  One thread per cell.
// ---------------------------------------------------------------------------
constexpr double kLimEps = 1.0e-20;
// The scalar a limiter launch advects: q - qoff with the step origin's q0,
// its tendency into dst plus the explicit-diffusion tendencies d6 and sm
// (theta: theta - 300 into tht; a moisture species: qoff 0, no diffusion).
struct LimScalar {
  const double* q;
  const double* q0;
  double qoff;
  double* dst;
  const double* d6;
  const double* sm;
};
__device__ inline double lim_pos(double a) { return a < 0.0 ? 0.0 : a; }
__device__ inline double lim_neg(double a) { return a > 0.0 ? 0.0 : a; }
__device__ inline double lim_min(double a, double b) { return b < a ? b : a; }
__device__ inline double lim_max(double a, double b) { return b > a ? b : a; }
__device__ inline double lim_upwind(double q_m, double q_p, double vel, double mu_face, double dxy, double dt) {
  const double cr = vel * dt / dxy / mu_face;
  const double acr = fabs(cr);
  const double w_m = 0.5 * lim_min(1.0, cr + acr);
  const double w_p = 0.5 * lim_max(-1.0, cr - acr);
  return mu_face * (dxy / dt) * (w_m * q_m + w_p * q_p);
}
__device__ inline double lim_scale(double flux, double room) {
  return flux > room ? lim_max(0.0, room / (flux + kLimEps)) : 1.0;
}
// Antidiffusive (a) and donor-cell (lo) flux at x face ii of row (k, j), or
// y face jj of column (k, i).
__device__ inline void lim_xface(const DGrid& g, const double* th, const double* th0, double qoff,
                                 const double* ru, const double* mut, int k, int j, int ii, double dxl, double dt,
                                 double& a, double& lo) {
  const int nx = g.nx;
  const int iw = ii == 0 ? nx - 1 : ii - 1;
  auto Q = [&](int d) { return th[im(g, k, j, wrapi(ii + d, nx))] - qoff; };
  const double muu = 0.5 * (mut[i2(g, j, ii)] + mut[i2(g, j, iw)]);
  const double mu_u = g.c1h[k] * muu + g.c2h[k];
  const double vel = ru[im(g, k, j, ii)];
  lo = lim_upwind(th0[im(g, k, j, iw)] - qoff, th0[im(g, k, j, ii)] - qoff, vel, mu_u, dxl, dt);
  a = vel * flux5(Q(-3), Q(-2), Q(-1), Q(0), Q(1), Q(2), vel) - lo;
}
__device__ inline void lim_yface(const DGrid& g, const double* th, const double* th0, double qoff,
                                 const double* rv, const double* mut, int k, int jj, int i, double dyl, double dt,
                                 double& a, double& lo) {
  const int ny = g.ny;
  const int js = jj == 0 ? ny - 1 : jj - 1;
  auto Q = [&](int d) { return th[im(g, k, wrapi(jj + d, ny), i)] - qoff; };
  const double muv = 0.5 * (mut[i2(g, jj, i)] + mut[i2(g, js, i)]);
  const double mu_v = g.c1h[k] * muv + g.c2h[k];
  const double vel = rv[im(g, k, jj, i)];
  lo = lim_upwind(th0[im(g, k, js, i)] - qoff, th0[im(g, k, jj, i)] - qoff, vel, mu_v, dyl, dt);
  a = vel * flux5(Q(-3), Q(-2), Q(-1), Q(0), Q(1), Q(2), vel) - lo;
}
// The same at z face f of column (j, i); faces 0 and nz carry no flux.
__device__ inline void lim_zface(const DGrid& g, const double* th, const double* th0, double qoff,
                                 const double* rom, const double* mut, int f, int j, int i, double dt, double& a,
                                 double& lo) {
  const int nz = g.nz;
  if (f == 0 || f == nz) {
    a = 0.0;
    lo = 0.0;
    return;
  }
  const size_t c = i2(g, j, i);
  const double mh = g.c1h[f] * mut[c] + g.c2h[f], mh1 = g.c1h[f - 1] * mut[c] + g.c2h[f - 1];
  const double dz = 2.0 / (g.rdnw[f] + g.rdnw[f - 1]);
  const double mu_face = 0.5 * (mh + mh1);
  const double r = rom[im(g, f, j, i)];
  const double cr = r * dt / dz / mu_face;
  const double acr = fabs(cr);
  const double w_m = 0.5 * lim_min(1.0, cr + acr), w_p = 0.5 * lim_max(-1.0, cr - acr);
  lo = mu_face * (dz / dt) * (w_m * (th0[im(g, f - 1, j, i)] - qoff) + w_p * (th0[im(g, f, j, i)] - qoff));
  auto Q = [&](int kk) { return th[im(g, kk, j, i)] - qoff; };
  double hi;
  if (f == 1)
    hi = r * (g.fnm[1] * Q(1) + g.fnp[1] * Q(0));
  else if (f == nz - 1)
    hi = r * (g.fnm[nz - 1] * Q(nz - 1) + g.fnp[nz - 1] * Q(nz - 2));
  else
    hi = r * flux3(Q(f - 2), Q(f - 1), Q(f), Q(f + 1), -r);
  a = hi - lo;
}
struct LimFaces {
  double ax0, lx0, ax1, lx1, ay0, ly0, ay1, ly1, az0, lz0, az1, lz1;
};
// Each cell's minus faces (x face i, y face j, z face k; and face nz above
// the top cell), computed once and stored.
__global__ void __launch_bounds__(256) k_lim_faces(DGrid g, DState s, DWork w, LimScalar a, double dt) {
  const int i = blockIdx.x * kFieldBx + threadIdx.x, j = blockIdx.y * kFieldBy + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const double dxl = 1.0 / (1.0 / g.dx), dyl = 1.0 / (1.0 / g.dy);
  const size_t c = im(g, k, j, i);
  lim_xface(g, a.q, a.q0, a.qoff, w.ru, s.mu_total, k, j, i, dxl, dt, w.lim_ax[c], w.lim_lx[c]);
  lim_yface(g, a.q, a.q0, a.qoff, w.rv, s.mu_total, k, j, i, dyl, dt, w.lim_ay[c], w.lim_ly[c]);
  lim_zface(g, a.q, a.q0, a.qoff, w.rom, s.mu_total, k, j, i, dt, w.lim_az[c], w.lim_lz[c]);
  if (k == g.nz - 1) {
    const size_t top = im(g, g.nz, j, i);
    lim_zface(g, a.q, a.q0, a.qoff, w.rom, s.mu_total, g.nz, j, i, dt, w.lim_az[top], w.lim_lz[top]);
  }
}
// The six faces of cell (k, j, i) from the stored face fluxes.
__device__ inline LimFaces lim_faces(const DGrid& g, const DWork& w, int k, int j, int i) {
  const int ip = i + 1 == g.nx ? 0 : i + 1, jp = j + 1 == g.ny ? 0 : j + 1;
  const size_t c = im(g, k, j, i), cx = im(g, k, j, ip), cy = im(g, k, jp, i), cz = im(g, k + 1, j, i);
  LimFaces F;
  F.ax0 = w.lim_ax[c];
  F.lx0 = w.lim_lx[c];
  F.ax1 = w.lim_ax[cx];
  F.lx1 = w.lim_lx[cx];
  F.ay0 = w.lim_ay[c];
  F.ly0 = w.lim_ly[c];
  F.ay1 = w.lim_ay[cy];
  F.ly1 = w.lim_ly[cy];
  F.az0 = w.lim_az[c];
  F.lz0 = w.lim_lz[c];
  F.az1 = w.lim_az[cz];
  F.lz1 = w.lim_lz[cz];
  return F;
}

__global__ void __launch_bounds__(256) k_lim_scale(DGrid g, DWork w, LimScalar a, int opt, double dt) {
  const int i = blockIdx.x * kFieldBx + threadIdx.x, j = blockIdx.y * kFieldBy + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
  const LimFaces F = lim_faces(g, w, k, j, i);
  const size_t c = im(g, k, j, i), c2 = i2(g, j, i);
  const double msf = g.msftx[c2];
  const double mass_old = g.c1h[k] * w.mut0[c2] + g.c2h[k];
  const double qoff = a.qoff;
  const double* th0 = a.q0;
  const double qo = th0[c] - qoff;
  const double div_low =
      msf * msf * (rdx * (F.lx1 - F.lx0) + rdy * (F.ly1 - F.ly0)) + msf * g.rdnw[k] * (F.lz1 - F.lz0);
  const double ph_low = mass_old * qo - dt * div_low;
  const double flux_out = dt * (msf * msf * (rdx * (lim_pos(F.ax1) - lim_neg(F.ax0)) +
                                             rdy * (lim_pos(F.ay1) - lim_neg(F.ay0))) +
                                msf * g.rdnw[k] * (lim_neg(F.az1) - lim_pos(F.az0)));
  if (opt == 1) {
    w.lim_sa[c] = lim_scale(flux_out, ph_low);
    return;
  }
  const int ip = i + 1 == nx ? 0 : i + 1, iw = i == 0 ? nx - 1 : i - 1;
  const int jp = j + 1 == ny ? 0 : j + 1, js = j == 0 ? ny - 1 : j - 1;
  const double nb[6] = {th0[im(g, k, j, iw)] - qoff, th0[im(g, k, j, ip)] - qoff,
                        th0[im(g, k, js, i)] - qoff, th0[im(g, k, jp, i)] - qoff,
                        th0[im(g, k > 0 ? k - 1 : 0, j, i)] - qoff,
                        th0[im(g, k < nz - 1 ? k + 1 : nz - 1, j, i)] - qoff};
  double qmax = qo, qmin = qo;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    qmax = lim_max(qmax, nb[n]);
    qmin = lim_min(qmin, nb[n]);
  }
  const double flux_in = -dt * (msf * msf * (rdx * (lim_neg(F.ax1) - lim_pos(F.ax0)) +
                                             rdy * (lim_neg(F.ay1) - lim_pos(F.ay0))) +
                                msf * g.rdnw[k] * (lim_pos(F.az1) - lim_neg(F.az0)));
  w.lim_sa[c] = lim_scale(flux_in, qmax * mass_old - ph_low);
  w.lim_sb[c] = lim_scale(flux_out, ph_low - qmin * mass_old);
}

// A horizontal face's limited antidiffusive flux a, m the cell on its minus
// side and p on its plus side; at a z face, m is the cell below, p above.
__device__ inline double lim_hface(int opt, double a, const double* sa, const double* sb, size_t m, size_t p) {
  if (opt == 1) return a > 0.0 ? sa[m] * a : (a < 0.0 ? sa[p] * a : a);
  return a > 0.0 ? lim_min(sa[p], sb[m]) * a : lim_min(sb[p], sa[m]) * a;
}
__device__ inline double lim_vface(int opt, double a, const double* sa, const double* sb, size_t below,
                                   size_t above) {
  if (opt == 1) return a < 0.0 ? sa[below] * a : (a > 0.0 ? sa[above] * a : a);
  return a < 0.0 ? lim_min(sa[above], sb[below]) * a : lim_min(sb[above], sa[below]) * a;
}

template <class TT>
__global__ void __launch_bounds__(256) k_lim_tend(DGrid g, DWork w, LimScalar a, int opt) {
  const int i = blockIdx.x * kFieldBx + threadIdx.x, j = blockIdx.y * kFieldBy + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
  const LimFaces F = lim_faces(g, w, k, j, i);
  const int ip = i + 1 == nx ? 0 : i + 1, iw = i == 0 ? nx - 1 : i - 1;
  const int jp = j + 1 == ny ? 0 : j + 1, js = j == 0 ? ny - 1 : j - 1;
  const size_t c = im(g, k, j, i);
  const double* sa = w.lim_sa;
  const double* sb = w.lim_sb;
  const double fx0 = lim_hface(opt, F.ax0, sa, sb, im(g, k, j, iw), c) + F.lx0;
  const double fx1 = lim_hface(opt, F.ax1, sa, sb, c, im(g, k, j, ip)) + F.lx1;
  const double fy0 = lim_hface(opt, F.ay0, sa, sb, im(g, k, js, i), c) + F.ly0;
  const double fy1 = lim_hface(opt, F.ay1, sa, sb, c, im(g, k, jp, i)) + F.ly1;
  const double fz0 = lim_vface(opt, F.az0, sa, sb, im(g, k > 0 ? k - 1 : 0, j, i), c) + F.lz0;
  const double fz1 = lim_vface(opt, F.az1, sa, sb, c, im(g, k + 1 < nz ? k + 1 : nz - 1, j, i)) + F.lz1;
  const double msf = g.msftx[i2(g, j, i)];
  double t = -msf * rdx * (fx1 - fx0);
  t = t - msf * rdy * (fy1 - fy0);
  t = t - g.rdnw[k] * (fz1 - fz0);
  const unsigned off = (unsigned)c * 8u;
  stt<TT>(a.dst, off, add_diff(t, a.d6, a.sm, off));
}

// ---------------------------------------------------------------------------
// large_step_coriolis (core/rk_addtend_dry.py:364-483), added to the coupled
// u and v tendencies after the large-step PGF (operational_mode.py:3052-3058),
// periodic run (no specified-boundary edge mask).  One thread per u-face
// column (blocks [0, nbu)) or v-face column (the rest), marching k:
//   u face: ut += (msfux/msfuy) f_u rv_quad - (e_u cosa_u) rw_quad
//   v face: vt += (-(msfvy/msfvx) f_v) ru_quad + ((msfvy/msfvx) e_v sina_v) rw_quad
// rv_quad / ru_quad average the coupled momenta of the four surrounding
// faces, rw_quad the coupled w of the two mass columns at levels k and k+1;
// x and y neighbours are clamped at the array edges, and the face masses are
// the edge-padded face averages of the stage mu_total.  Only launched when
// the grid rotates (f or e nonzero).
// ---------------------------------------------------------------------------
template <class TT>
__device__ __forceinline__ void b_coriolis(const Blk& B, DGrid g, DState s, DWork w,
                                           unsigned nbu) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t pl = (size_t)ny * nx, upl = (size_t)ny * (nx + 1), vpl = (size_t)(ny + 1) * nx;
  const bool ufaces = B.x < nbu;
  const size_t t = (size_t)(ufaces ? B.x : B.x - nbu) * B.bdx + B.tx;
  if (ufaces ? t >= upl : t >= vpl) return;
  auto rw_at = [&](int k, size_t c) {
    return (s.w[(size_t)k * pl + c] * (g.c1f[k] * s.mu_total[c] + g.c2f[k])) / g.msfty[c];
  };
  if (ufaces) {
    const int j = (int)(t / (size_t)(nx + 1)), i = (int)(t % (size_t)(nx + 1));
    const int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
    const size_t cl = i2(g, j, il), cr = i2(g, j, ir);
    const double msf_u = g.msfux[t] / g.msfuy[t];
    const double f_u = 0.5 * (g.fcor[cl] + g.fcor[cr]);
    const double e_u = 0.5 * (g.ecor[cl] + g.ecor[cr]);
    const double cosa_u = 0.5 * (g.cosa[cl] + g.cosa[cr]);
    // the four v faces (rows j, j+1; columns il, ir), their face mu and msfvx
    size_t fv[4];
    double muv[4], mvx[4];
    for (int q = 0; q < 4; ++q) {
      const int jj = j + (q >> 1), ii = (q & 1) ? ir : il;
      fv[q] = i2v(g, jj, ii);
      muv[q] = 0.5 * (s.mu_total[i2(g, clampi(jj - 1, 0, ny - 1), ii)] + s.mu_total[i2(g, clampi(jj, 0, ny - 1), ii)]);
      mvx[q] = g.msfvx[fv[q]];
    }
    double rwl = rw_at(0, cl), rwr = rw_at(0, cr);
    for (int k = 0; k < nz; ++k) {
      double rv[4];
      for (int q = 0; q < 4; ++q) rv[q] = (s.v[(size_t)k * vpl + fv[q]] * (g.c1h[k] * muv[q] + g.c2h[k])) / mvx[q];
      const double rv_quad = 0.25 * (((rv[0] + rv[1]) + rv[2]) + rv[3]);
      const double rwl1 = rw_at(k + 1, cl), rwr1 = rw_at(k + 1, cr);
      const double rw_quad = 0.25 * (((rwl + rwl1) + rwr) + rwr1);
      double cor = (msf_u * f_u) * rv_quad;
      cor = cor - (e_u * cosa_u) * rw_quad;
      const unsigned off = (unsigned)((size_t)k * upl + t) * 8u;
      stt<TT>(w.ut, off, ldt<TT>(w.ut, off) + cor);
      rwl = rwl1;
      rwr = rwr1;
    }
  } else {
    const int j = (int)(t / (size_t)nx), i = (int)(t % (size_t)nx);
    const int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
    const size_t cs = i2(g, js, i), cn = i2(g, jn, i);
    const double msf_v = g.msfvy[t] / g.msfvx[t];
    const double f_v = 0.5 * (g.fcor[cs] + g.fcor[cn]);
    const double e_v = 0.5 * (g.ecor[cs] + g.ecor[cn]);
    const double sina_v = 0.5 * (g.sina[cs] + g.sina[cn]);
    // the four u faces (rows js, jn; columns i, i+1), their face mu and msfuy
    size_t fu[4];
    double muu[4], muy[4];
    for (int q = 0; q < 4; ++q) {
      const int jj = (q >> 1) ? jn : js, ii = i + (q & 1);
      fu[q] = i2u(g, jj, ii);
      muu[q] = 0.5 * (s.mu_total[i2(g, jj, clampi(ii - 1, 0, nx - 1))] + s.mu_total[i2(g, jj, clampi(ii, 0, nx - 1))]);
      muy[q] = g.msfuy[fu[q]];
    }
    double rws = rw_at(0, cs), rwn = rw_at(0, cn);
    for (int k = 0; k < nz; ++k) {
      double ru[4];
      for (int q = 0; q < 4; ++q) ru[q] = (s.u[(size_t)k * upl + fu[q]] * (g.c1h[k] * muu[q] + g.c2h[k])) / muy[q];
      const double ru_quad = 0.25 * (((ru[0] + ru[1]) + ru[2]) + ru[3]);
      const double rws1 = rw_at(k + 1, cs), rwn1 = rw_at(k + 1, cn);
      const double rw_quad = 0.25 * (((rws + rws1) + rwn) + rwn1);
      double cor = ((-msf_v) * f_v) * ru_quad;
      cor = cor + ((msf_v * e_v) * sina_v) * rw_quad;
      const unsigned off = (unsigned)((size_t)k * vpl + t) * 8u;
      stt<TT>(w.vt, off, ldt<TT>(w.vt, off) + cor);
      rws = rws1;
      rwn = rwn1;
    }
  }
}

template <class TT>
__global__ void __launch_bounds__(256) k_coriolis(DGrid g, DState s, DWork w, unsigned nbu) {
  b_coriolis<TT>(hw_blk(), g, s, w, nbu);
}

// Launch grids of the stage kernels, which the persistent ensemble step
// gives its virtual blocks too.  Blocks are 256 threads (64 x 4 for the
// tiles) except k_rom's 64.
inline unsigned level_chunks(const DGrid& g) { return (unsigned)((g.nz + kChunk - 1) / kChunk); }
inline unsigned rom_blocks(const DGrid& g) { return (unsigned)((g.nx + 63) / 64) * (unsigned)g.ny; }
inline unsigned tend_u_tiles(const DGrid& g) {
  return (unsigned)((g.nx + 1 + 61) / 62) * (unsigned)((g.ny + kTendURows - 1) / kTendURows);
}
inline unsigned tend_v_tiles(const DGrid& g) {
  return (unsigned)((g.nx + 61) / 62) * (unsigned)((g.ny + 1 + kTendURows - 1) / kTendURows);
}
inline unsigned tend_t_tiles(const DGrid& g) {
  return (unsigned)((g.nx + 61) / 62) * (unsigned)((g.ny + kTendURows - 1) / kTendURows);
}
inline unsigned col_setup_tiles(const DGrid& g) {
  return (unsigned)((g.nx + kColOwned - 1) / kColOwned) * (unsigned)((g.ny + kColRows - 1) / kColRows);
}
inline unsigned coriolis_u_blocks(const DGrid& g) { return blocks_for((size_t)g.ny * (g.nx + 1), 256); }
inline unsigned coriolis_blocks(const DGrid& g) {
  return coriolis_u_blocks(g) + blocks_for((size_t)(g.ny + 1) * g.nx, 256);
}

template <class TT>
void launch_coriolis(const DGrid& g, const DState& s, const DWork& w, hipStream_t st) {
  k_coriolis<TT><<<coriolis_blocks(g), 256, 0, st>>>(g, s, w, coriolis_u_blocks(g));
}

inline dim3 cell_grid(const DGrid& g) {
  return dim3((unsigned)((g.nx + kFieldBx - 1) / kFieldBx), (unsigned)((g.ny + kFieldBy - 1) / kFieldBy),
              (unsigned)g.nz);
}

// One limited advection (opt 1 or 2) of the scalar `a`; the face and scale
// arrays are shared, so limiter launches on one stream run one after another.
template <class TT>
void launch_limiter(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w, const LimScalar& a,
                    int opt, hipStream_t st) {
  const dim3 bs(kFieldBx, kFieldBy);
  const dim3 grid = cell_grid(g);
  k_lim_faces<<<grid, bs, 0, st>>>(g, s, w, a, nl.dt);
  k_lim_scale<<<grid, bs, 0, st>>>(g, w, a, opt, nl.dt);
  k_lim_tend<TT><<<grid, bs, 0, st>>>(g, w, a, opt);
}

// ---------------------------------------------------------------------------
// The moisture.  moisture_coupling_factors (dynamics/acoustic_wrf.py:303-319)
// and moist_cqw_calc_face (core/advance_w.py:195-219) of the stage-entry
// species, one thread per mass cell: qtot = qv + qc + qr + qi + qs + qg;
// cqu = 1/(1 + (qtot_west + qtot_east)/2) on the cell's west u face (and the
// last face of the row, the neighbour clamped at the edge), cqv likewise on
// the south v face; the raw cqw = (qtot(k) + qtot(k-1))/2 on w face k, 0 at
// the surface and the lid.  k_tend_u/v scale the large-step PGF by cqu/cqv,
// k_uv the acoustic one, k_col_setup forms rw_tend and the implicit-w
// coefficients with cq1 = 1/(1 + cqw) and cq2 = cqw cq1, and k_column_w the
// explicit part of the w equation with cq1.
// ---------------------------------------------------------------------------
__device__ inline double qtot_at(const DState& s, size_t c) {
  return ((((s.q[0][c] + s.q[1][c]) + s.q[2][c]) + s.q[3][c]) + s.q[4][c]) + s.q[5][c];
}
__global__ void __launch_bounds__(256) k_moist_factors(DGrid g, DState s, DWork w) {
  const int i = blockIdx.x * kFieldBx + threadIdx.x, j = blockIdx.y * kFieldBy + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double qc = qtot_at(s, im(g, k, j, i));
  const double qw = i > 0 ? qtot_at(s, im(g, k, j, i - 1)) : qc;
  const double qs = j > 0 ? qtot_at(s, im(g, k, j - 1, i)) : qc;
  w.cqu[iu(g, k, j, i)] = 1.0 / (1.0 + 0.5 * (qw + qc));
  if (i == nx - 1) w.cqu[iu(g, k, j, nx)] = 1.0 / (1.0 + 0.5 * (qc + qc));
  w.cqv[iv(g, k, j, i)] = 1.0 / (1.0 + 0.5 * (qs + qc));
  if (j == ny - 1) w.cqv[iv(g, k, ny, i)] = 1.0 / (1.0 + 0.5 * (qc + qc));
  w.cqwr[im(g, k, j, i)] = k > 0 ? 0.5 * (qc + qtot_at(s, im(g, k - 1, j, i))) : 0.0;
  if (k == nz - 1) w.cqwr[im(g, nz, j, i)] = 0.0;
}

// _apply_moisture_large_step (runtime/operational_mode.py:3273-3316) after a
// stage's acoustic loop: q = (m_origin q_origin + dt_rk qt) / m_new with the
// column masses c1h mu + c2h of the step origin and of the finished stage,
// every species of the cell in one thread.
__global__ void __launch_bounds__(256) k_moist_update(DGrid g, DState s, DWork w, double dt_rk) {
  const int i = blockIdx.x * kFieldBx + threadIdx.x, j = blockIdx.y * kFieldBy + threadIdx.y, k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const size_t c2 = i2(g, j, i), c = im(g, k, j, i);
  const double mass_old = g.c1h[k] * w.mut0[c2] + g.c2h[k];
  const double mass_new = g.c1h[k] * s.mu_total[c2] + g.c2h[k];
  const double inv_mass_new = 1.0 / mass_new;
#pragma unroll
  for (int sp = 0; sp < kNumMoist; ++sp) s.q[sp][c] = (mass_old * w.q0[sp][c] + dt_rk * w.qt[sp][c]) * inv_mass_new;
}

hipError_t launch_moist_update(const DGrid& g, const DState& s, const DWork& w, double dt_rk, hipStream_t st,
                               Prof* prof) {
  WD_TIMED(prof, KID_MOIST_UPDATE, st,
           (k_moist_update<<<cell_grid(g), dim3(kFieldBx, kFieldBy), 0, st>>>(g, s, w, dt_rk)));
  return hipGetLastError();
}

hipError_t launch_stage_setup(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              int rk_step, double dts, hipStream_t st, Prof* prof) {
  const unsigned bs = 256;
  const size_t C = (size_t)g.ny * g.nx;
  const size_t F = (size_t)g.ny * (g.nx + 1) + (size_t)(g.ny + 1) * g.nx;
  const unsigned kchunks = level_chunks(g);
  if (w.d6u) WD_TIMED(prof, KID_DIFF6, st, (launch_diff6(g, nl, s, w, st)));
  if (w.xkmh) WD_TIMED(prof, KID_SMAG, st, (launch_smag(g, nl, s, w, st)));
  if (w.ck_dz) WD_TIMED(prof, KID_CONSTK, st, (launch_constk(g, nl, s, w, st)));
  // theta's limited advection replaces the plain one on the final stage
  const bool limit_theta = nl.scalar_adv_opt != 0 && rk_step == 3;
  const LimScalar theta_lim{s.theta, w.th0, kT0, w.tht, w.d6t, w.smt};
  // the moisture factors of the stage-entry species (k_tend_u/v read cqu/cqv)
  if (nl.moist)
    WD_TIMED(prof, KID_MOIST, st,
             (k_moist_factors<<<cell_grid(g), dim3(kFieldBx, kFieldBy), 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_RUV, st,
           (k_ruv<<<dim3(blocks_for(C, bs), kchunks), bs, 0, st>>>(g, s.u, s.v, s.mu_total, w.ru, w.rv)));
  WD_TIMED(prof, KID_ROM, st, (k_rom<<<rom_blocks(g), 64, 0, st>>>(g, w.ru, w.rv, w.rom)));
  WD_TIMED(prof, KID_PREP_2D, st, (k_prep_2d<<<blocks_for(C, bs), bs, 0, st>>>(g, s, w, rk_step)));
  // k_prep_col writes the large-step diagnostics the tendency kernels read
  WD_TIMED(prof, KID_PREP_COL, st, (k_prep_col<<<blocks_for(C, bs), bs, 0, st>>>(g, s, w)));
  const size_t CU = (size_t)g.ny * (g.nx + 1), CV = (size_t)(g.ny + 1) * g.nx;
  const unsigned tu_tiles = tend_u_tiles(g), tv_tiles = tend_v_tiles(g), tt_tiles = tend_t_tiles(g);
  const dim3 tu_block(64, kTendURows);
  if (nl.mixed) {
    WD_TIMED(prof, KID_TEND_U, st, (k_tend_u<float, false><<<tu_tiles, tu_block, 0, st>>>(g, s, w)));
    WD_TIMED(prof, KID_TEND_V, st, (k_tend_v<float, false><<<tv_tiles, tu_block, 0, st>>>(g, s, w)));
    WD_TIMED(prof, KID_TEND_W, st, (k_tend_w<float><<<blocks_for(C, bs), bs, 0, st>>>(g, s, w)));
    if (limit_theta)
      WD_TIMED(prof, KID_LIMITER, st, (launch_limiter<float>(g, nl, s, w, theta_lim, nl.scalar_adv_opt, st)));
    else
      WD_TIMED(prof, KID_TEND_T, st,
               (k_tend_t<float><<<tt_tiles, tu_block, 0, st>>>(g, w, s.theta, kT0, w.tht, w.d6t, w.smt)));
  } else {
    if (w.cqu) {
      WD_TIMED(prof, KID_TEND_U, st, (k_tend_u<double, true><<<tu_tiles, tu_block, 0, st>>>(g, s, w)));
      WD_TIMED(prof, KID_TEND_V, st, (k_tend_v<double, true><<<tv_tiles, tu_block, 0, st>>>(g, s, w)));
    } else {
      WD_TIMED(prof, KID_TEND_U, st, (k_tend_u<double, false><<<tu_tiles, tu_block, 0, st>>>(g, s, w)));
      WD_TIMED(prof, KID_TEND_V, st, (k_tend_v<double, false><<<tv_tiles, tu_block, 0, st>>>(g, s, w)));
    }
    WD_TIMED(prof, KID_TEND_W, st, (k_tend_w<double><<<blocks_for(C, bs), bs, 0, st>>>(g, s, w)));
    if (limit_theta)
      WD_TIMED(prof, KID_LIMITER, st, (launch_limiter<double>(g, nl, s, w, theta_lim, nl.scalar_adv_opt, st)));
    else
      WD_TIMED(prof, KID_TEND_T, st,
               (k_tend_t<double><<<tt_tiles, tu_block, 0, st>>>(g, w, s.theta, kT0, w.tht, w.d6t, w.smt)));
  }
  // the species' coupled tendencies with the same transport velocities: the
  // plain h5/v3 advection, the limiter on the final stage from the step
  // origin (operational_mode.py:3162-3230); FP64 form only
  if (nl.moist && nl.moist_adv_opt != 0)
    for (int sp = 0; sp < kNumMoist; ++sp) {
      if (rk_step == 3)
        WD_TIMED(prof, KID_MOIST, st,
                 (launch_limiter<double>(g, nl, s, w, LimScalar{s.q[sp], w.q0[sp], 0.0, w.qt[sp], nullptr, nullptr},
                                         nl.moist_adv_opt, st)));
      else
        WD_TIMED(prof, KID_MOIST, st,
                 (k_tend_t<double><<<tt_tiles, tu_block, 0, st>>>(g, w, s.q[sp], 0.0, w.qt[sp], nullptr, nullptr)));
    }
  // the Coriolis force joins the coupled u/v tendencies after the PGF
  if (g.rot) {
    if (nl.mixed)
      WD_TIMED(prof, KID_CORIOLIS, st, (launch_coriolis<float>(g, s, w, st)));
    else
      WD_TIMED(prof, KID_CORIOLIS, st, (launch_coriolis<double>(g, s, w, st)));
  }
  WD_TIMED(prof, KID_PREP_FACES, st, (k_prep_faces<<<blocks_for(F, bs), bs, 0, st>>>(g, w)));
  WD_TIMED(prof, KID_PREP_U, st, (k_prep_u<<<dim3(blocks_for(CU, bs), kchunks), bs, 0, st>>>(g, s, w)));
  WD_TIMED(prof, KID_PREP_V, st, (k_prep_v<<<dim3(blocks_for(CV, bs), kchunks), bs, 0, st>>>(g, s, w)));
  const unsigned col_tiles = col_setup_tiles(g);
  const dim3 col_block(64, kColRows);
  if (nl.mixed) {
    WD_TIMED(prof, KID_COL_SETUP, st, (k_col_setup<float, false><<<col_tiles, col_block, 0, st>>>(g, nl, s, w, dts)));
  } else if (w.cqwr) {
    WD_TIMED(prof, KID_COL_SETUP, st, (k_col_setup<double, true><<<col_tiles, col_block, 0, st>>>(g, nl, s, w, dts)));
  } else {
    WD_TIMED(prof, KID_COL_SETUP, st, (k_col_setup<double, false><<<col_tiles, col_block, 0, st>>>(g, nl, s, w, dts)));
  }
  return hipGetLastError();
}

}  // namespace wd
