// Shared device definitions for the wrfdyn HIP kernels (gfx950, FP64).
//
// Physical constants follow the reference dycore:
//   src/gpuwrf/dynamics/acoustic_wrf.py:27-75 (R_D, CP_D=7R/2, P0, g=9.81,
//   MIN_ACOUSTIC_ALT=0.1), core/advance_w.py:31 (g), core/acoustic.py:845-846
//   (emdiv=0.01, smdiv=0.1), runtime/operational_mode.py:202 (1/6370e3).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "wrfdyn_math.h"

namespace wd {

constexpr double kRd = 287.0;
constexpr double kCp = 7.0 * kRd / 2.0;
constexpr double kP0 = 100000.0;
constexpr double kG = 9.81;
constexpr double kT0 = 300.0;
constexpr double kEmdiv = 0.01;
constexpr double kSmdiv = 0.1;
constexpr double kMinAcousticAlt = 0.1;
constexpr double kMinAlt = 1.0e-8;
constexpr double kMinPressure = 1.0;
constexpr double kWAlpha = 0.3;
constexpr double kWBeta = 1.0;
constexpr double kCpovcv = kCp / (kCp - kRd);
constexpr double kCvpm = -(kCp - kRd) / kCp;
constexpr double kReradius = 1.0 / 6370.0e3;
constexpr double kPi = 3.141592653589793;

// Grid + metrics as passed to every kernel (by value, kernarg segment).
struct DGrid {
  int nx, ny, nz;
  double dx, dy, rdx, rdy;
  const double *c1h, *c2h, *c1f, *c2f;
  const double *dnw, *rdnw, *dn, *rdn, *fnm, *fnp;
  double cf1, cf2, cf3, cfn, cfn1;
  const double *msftx, *msfty, *msfux, *msfuy, *msfvx, *msfvy;
  const double *c3h, *c4h, *c3f, *c4f;  // dry-pressure coefficients (hypsometric_opt 2)
  double p_top;
  // Coriolis parameters on mass points (null unless rot): f, e, sina, cosa
  const double *fcor, *ecor, *sina, *cosa;
  int rot;              // 1: f or e nonzero somewhere, the Coriolis terms run
  int hypsometric_opt;  // 2: log-pressure al (WRF calc_p_rho_phi), else linear
  int xcd_remap;  // 1: walk tiles XCD-contiguously (xcd_tile), 0: hardware order
};

struct DNamelist {
  double dt;
  int ns;
  double epssm;
  int top_lid, w_damping, damp_opt;
  double dampcoef, zdamp;
  int diff_6th_opt;        // 0 off, 1/2 the 6th-order filter (monotonic)
  double diff_6th_factor;
  int smag;                // 1: 2-D Smagorinsky horizontal diffusion (diff_opt 1, km_opt 4)
  double c_s;
  int scalar_adv_opt;      // theta on the final RK stage: 0 plain, 1 positive-definite, 2 monotonic
  double const_nu;         // > 0: constant-K diffusion with this K (m^2/s)
  int moist;               // 1: the state carries the six moisture species
  int moist_adv_opt;       // with moist: 0 no transport, 1/2 transport, limited on the final stage
  int mixed;  // 1: mixed-precision form (FP32 tendencies and flux interpolants), 0: plain FP64
};

// Caller state (device pointers).
constexpr int kNumMoist = 6;  // qv, qc, qr, qi, qs, qg
struct DState {
  double *u, *v, *w, *theta, *p_total, *p_pert, *ph_total, *ph_pert, *mu_total, *mu_pert, *mudf;
  double* q[kNumMoist];  // the moisture species (all null for a dry state)
};

// Per-handle device workspace.
struct DWork {
  // RK1 reference (step origin) copies
  double *u0, *v0, *w0, *th0, *ph0, *mu0;
  // stage transport velocities
  double *ru, *rv, *rom;
  // large-step tendencies
  double *ut, *vt, *wt, *tht;
  // stage diagnostics (large-step PGF)
  double *alt_raw, *al_ls, *php_ls;
  // prep (2-D)
  double *mub, *mut, *muts0, *mu_work0, *mu, *muts, *muave;
  double *muu, *muus, *muv, *muvs;
  // prep (3-D)
  double *u_work, *v_work, *w_work, *th_work, *ph_work;
  double *c2a, *alt, *pb, *ph1t, *php;  // ph1t = origin ph' + base ph (stage constant)
  double *p, *al, *pm1;
  // stage column setup
  double *rw_tend, *ph_tend, *ca, *calpha, *cgamma;
  // acoustic carried (ww: only levels 0 and nz carry between substeps)
  double *ww, *t2ave;
  // column-march scratch: phi-equation rhs and the Thomas forward sweep
  // (k_column_w); rhs also carries dvdxi between k_column_mu's two sweeps
  double *rhs, *wfwd;
  // the 6th-order filter's coupled tendencies of u, v, w, theta (null unless
  // diff_6th_opt != 0), added to the RK tendencies before the PGF
  double *d6u, *d6v, *d6w, *d6t;
  // the 2-D Smagorinsky diffusion: the eddy viscosity on mass cells and the
  // coupled tendencies of u, v, w, theta (null unless diff_opt 1, km_opt 4),
  // added after the filter's
  double *xkmh, *smu, *smv, *smw, *smt;
  // the limited theta advection (null unless scalar_adv_opt != 0): the step
  // origin's mu_total, the per-cell scales (positive-definite: lim_sa;
  // monotonic: inflow lim_sa, outflow lim_sb) and
  double *mut0, *lim_sa, *lim_sb;
  // its face fluxes: antidiffusive (a) and donor-cell (l) at each cell's
  // x, y and z minus face (z: nz + 1 faces)
  double *lim_ax, *lim_lx, *lim_ay, *lim_ly, *lim_az, *lim_lz;
  // the constant-K diffusion (null unless const_nu > 0; its tendencies go to
  // smu..smt): each column's summed layer thickness and the grid mean dz
  double *ck_col, *ck_dz;
  // the moisture (null for a dry state): the coupling factors of the stage
  // species on u faces, v faces and w faces (the raw calc_cq field, 0 at the
  // surface and the lid), and with transport the step-origin species and the
  // stage's coupled tendencies d(mu q)/dt
  double *cqu, *cqv, *cqwr;
  double* q0[kNumMoist];
  double* qt[kNumMoist];
};

__host__ __device__ inline size_t im(const DGrid& g, int k, int j, int i) {
  return ((size_t)k * g.ny + (size_t)j) * g.nx + (size_t)i;
}
__host__ __device__ inline size_t iu(const DGrid& g, int k, int j, int i) {
  return ((size_t)k * g.ny + (size_t)j) * (g.nx + 1) + (size_t)i;
}
__host__ __device__ inline size_t iv(const DGrid& g, int k, int j, int i) {
  return ((size_t)k * (g.ny + 1) + (size_t)j) * g.nx + (size_t)i;
}
__host__ __device__ inline size_t i2(const DGrid& g, int j, int i) { return (size_t)j * g.nx + (size_t)i; }
__host__ __device__ inline size_t i2u(const DGrid& g, int j, int i) { return (size_t)j * (g.nx + 1) + (size_t)i; }
__host__ __device__ inline size_t i2v(const DGrid& g, int j, int i) { return (size_t)j * g.nx + (size_t)i; }

__device__ inline int wrapi(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}
// gfx950 dispatches the workgroups of a launch round-robin over its 8 XCDs
// (linear workgroup id b runs on XCD b % 8), and each XCD has its own L2.
// xcd_tile maps b to a logical tile id so that every XCD walks one contiguous
// run of the n tiles: neighbouring tiles, which share halo rows and the cache
// lines split between them, then meet in the same L2.  A bijection on [0, n).
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_tile(unsigned b, unsigned n) {
  const unsigned x = b % kNumXcd, s = b / kNumXcd, q = n / kNumXcd, r = n % kNumXcd;
  return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}
// The block a kernel body runs as: its place in the launch grid (x, y, and
// the grid's width gx), the thread's place in the block (tx, ty) and the
// block's width bdx.  A __global__ kernel hands its body the hardware's
// (hw_blk); the persistent ensemble step hands it the virtual block it gave
// to one of its resident workgroups, so the same body serves both.
struct Blk {
  unsigned x, y, gx;
  unsigned tx, ty, bdx;
};
__device__ __forceinline__ Blk hw_blk() {
  return Blk{blockIdx.x, blockIdx.y, gridDim.x, threadIdx.x, threadIdx.y, blockDim.x};
}
__device__ inline unsigned tile_id(const DGrid& g, const Blk& b) {
  return g.xcd_remap ? xcd_tile(b.x, b.gx) : b.x;
}
__device__ inline unsigned tile_id(const DGrid& g) { return tile_id(g, hw_blk()); }

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Loads and stores at a 32-bit byte offset from a wave-uniform base pointer.
// The compiler then addresses with the base in SGPRs plus one 32-bit VGPR
// offset (the global_load/store saddr form) instead of a 64-bit VGPR address
// per access, which halves address registers and address arithmetic in the
// k-marching kernels.  Every 3-D array must stay below 4 GiB (wd_create
// checks the grid).
__device__ __forceinline__ double ldo(const double* base, unsigned byte_off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ void sto(double* base, unsigned byte_off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// t plus the explicit-diffusion tendencies present at byte offset off: the
// 6th-order filter's, then the Smagorinsky diffusion's, in the order the
// reference adds them (runtime/operational_mode.py:2870-2963)
__device__ __forceinline__ double add_diff(double t, const double* d6, const double* sm, unsigned off) {
  if (d6) t = t + ldo(d6, off);
  if (sm) t = t + ldo(sm, off);
  return t;
}

// Tendency arrays are stored as T (double, or float in the mixed-precision
// form) in slots sized for doubles.  ldt/stt take the byte offset the element
// would have in a double array (off8 = 8 * index) and access element `index`
// of the T array, converting to and from double.
template <class T>
__device__ __forceinline__ double ldt(const double* base, unsigned off8) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return ldo(base, off8);
  } else {
    return (double)*reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (off8 >> 1));
  }
}
template <class T>
__device__ __forceinline__ void stt(double* base, unsigned off8, double v) {
  if constexpr (sizeof(T) == sizeof(double)) {
    sto(base, off8, v);
  } else {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (off8 >> 1)) = (float)v;
  }
}

// Cross-lane moves within a 64-lane wavefront (DPP, no memory access):
// kDppWaveShl1 gives each lane the value of lane + 1, kDppWaveShr1 that of
// lane - 1.  With bound_ctrl off, the lane whose source lies outside the wave
// (63, or 0) keeps `old`.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double dpp_move(double old, double src) {
  const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// jnp.sign: -1, +1, and the argument itself for +-0 and NaN.
__device__ inline double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
// jnp.where(|x| > floor, x, floor)
__device__ inline double safe_pos(double x, double floor) { return fabs(x) > floor ? x : floor; }

// _inverse_density_from_theta_pressure, dynamics/acoustic_wrf.py:178-195 (qvf=1).
// hypsometric_opt 2 (WRF calc_p_rho_phi log form, rk_addtend_dry.py:196-222,
// acoustic_wrf.py:394-427): the specific volume of layer k from its
// geopotential depth dph and the column mass, dph / phm / ln(pfd / pfu), with
// the face and half-level dry pressures c3 mass + c4 + p_top.
__device__ inline double log_alpha(const DGrid& g, int k, double dph, double mass) {
  const double pfu = g.c3f[k + 1] * mass + g.c4f[k + 1] + g.p_top;
  const double pfd = g.c3f[k] * mass + g.c4f[k] + g.p_top;
  const double phm = g.c3h[k] * mass + g.c4h[k] + g.p_top;
  return dph / phm / wd_log(pfd / pfu);
}
__device__ inline double alt_from_theta_p(double theta, double p) {
  double ps = p > kMinPressure ? p : kMinPressure;
  return (((kRd / kP0) * theta) * 1.0) * wd_pow(ps / kP0, kCvpm);
}
// _pressure_from_theta_alt, dynamics/acoustic_wrf.py:198-211.
__device__ inline double p_from_theta_alt(double theta, double alt) {
  double as = alt > kMinAlt ? alt : kMinAlt;
  double arg = ((kRd * theta) * 1.0) / (kP0 * as);
  double a2 = arg > 1.0e-12 ? arg : 1.0e-12;
  return kP0 * wd_pow(a2, kCpovcv);
}

// flux5_face_periodic, dynamics/flux_advection.py:53-88.
__device__ inline double flux5(double q_im3, double q_im2, double q_im1, double q_i, double q_ip1,
                               double q_ip2, double vel) {
  double f6 = ((37.0 * (q_i + q_im1) - 8.0 * (q_ip1 + q_im2)) + (q_ip2 + q_im3)) / 60.0;
  double corr = (((q_ip2 - q_im3) - 5.0 * (q_ip1 - q_im2)) + 10.0 * (q_i - q_im1)) / 60.0;
  return f6 - sgn(vel) * corr;
}
// flux3 with caller-chosen upwind velocity (flux_advection.py:537-547, 1388-1397, 1452-1467).
__device__ inline double flux3(double q_km2, double q_km1, double q_k, double q_kp1, double velz) {
  double f4 = (7.0 * (q_k + q_km1) - (q_kp1 + q_km2)) / 12.0;
  double corr = ((q_kp1 - q_km2) - 3.0 * (q_k - q_km1)) / 12.0;
  return f4 + sgn(velz) * corr;
}

// ---- mixed-precision flux interpolation ----
// The flux5 and flux3 interpolants are affine with weights summing to one,
// so interpolating q - qc and adding qc back is exact algebra.  The mixed
// form evaluates the interpolant of the deviations q - qc (qc: the value at
// the cell the divergence is taken for) in FP32 and keeps the mean part in
// FP64: for the pair of faces of a cell,
//   v1 F(q)_1 - v0 F(q)_0 = [v1 F(q-qc)_1 - v0 F(q-qc)_0]_FP32 + [qc (v1 - v0)]_FP64,
// so FP32 rounding acts on the small deviations only, and the transport
// divergence, which the mass flux nearly cancels, stays FP64.
__device__ inline float sgnf(double x) { return x > 0.0 ? 1.0f : (x < 0.0 ? -1.0f : (float)x); }
__device__ inline float flux5f(float q_im3, float q_im2, float q_im1, float q_i, float q_ip1, float q_ip2, float s) {
  float f6 = ((37.0f * (q_i + q_im1) - 8.0f * (q_ip1 + q_im2)) + (q_ip2 + q_im3)) / 60.0f;
  float corr = (((q_ip2 - q_im3) - 5.0f * (q_ip1 - q_im2)) + 10.0f * (q_i - q_im1)) / 60.0f;
  return f6 - s * corr;
}
__device__ inline float flux3f(float q_km2, float q_km1, float q_k, float q_kp1, float s) {
  float f4 = (7.0f * (q_k + q_km1) - (q_kp1 + q_km2)) / 12.0f;
  float corr = ((q_kp1 - q_km2) - 3.0f * (q_k - q_km1)) / 12.0f;
  return f4 + s * corr;
}
// v1 * flux5(x[1..6], v1) - v0 * flux5(x[0..5], v0) for the cell whose value
// is x[3]; qc is x[3] in the units the interpolated quantity uses.
__device__ inline double hdiv5_mixed(const double* x, double qc, double v0, double v1) {
  float d[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) d[k] = (float)(x[k] - x[3]);
  const float f0 = flux5f(d[0], d[1], d[2], d[3], d[4], d[5], sgnf(v0));
  const float f1 = flux5f(d[1], d[2], d[3], d[4], d[5], d[6], sgnf(v1));
  const float dev = (float)v1 * f1 - (float)v0 * f0;
  return (double)dev + qc * (v1 - v0);
}

// post-pg_buoy_w cqw with qtot=0: 0 at faces 0 and nz, 1 inside
// (core/advance_w.py:266-290).
__device__ inline double cqw_at(int kf, int nz) { return (kf >= 1 && kf < nz) ? 1.0 : 0.0; }

}  // namespace wd
