/*
 * CPU oracle: FP64 restatement of the reference's split-explicit ARW large
 * step (periodic idealized path, dry or with the six moisture species).
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP product path, and the
 * timed CPU baseline.  Never linked into or called by the product.
 *
 * Every function below cites the reference function (path:line relative to
 * the reference checkout) that it restates.  Expressions keep the reference's
 * left-to-right evaluation order, and the file is compiled with
 * -ffp-contract=off so no multiply-add is fused.  Where the reference leaves
 * an order to XLA (jnp.sum, jnp.cumsum over the vertical) the oracle sums
 * sequentially from k=0 upwards.
 *
 * Reference constants (src/gpuwrf/dynamics/acoustic_wrf.py:27-75,
 * core/advance_w.py:31, runtime/operational_mode.py:202):
 *   R_D=287, CP_D=7*R_D/2, P0=1e5, CPOVCV=CP_D/(CP_D-R_D),
 *   CVPM=-(CP_D-R_D)/CP_D, g=9.81, MIN_ACOUSTIC_ALT=0.1, re=6370e3,
 *   emdiv=0.01, smdiv=0.1 (core/acoustic.py:845-846), t0=300.
 *
 * How it is pinned (tests/test_oracle_golden.py, test_oracle_invariants.py,
 * test_oracle_terrain_pins.py, test_oracle_coriolis.py, test_oracle_straka.py):
 * built with the reference's v0.13 constants (ORC_LEGACY_V013) it reproduces
 * the reference-held warm-bubble golden scalars (2-D, 80x40, 5000 steps) to
 * <= 4.9e-7, which pins the whole x-z large step incl. advection, PGF, prep,
 * advance_uv/mu_t/w, Thomas, damp_opt=3, calc_p_rho and finish, and the
 * Straka density-current golden scalars (9000 steps) to <= 4e-4, which pins
 * the constant-K diffusion; the y path is pinned against the x path by x<->y
 * transposition; Thomas, calc_coef_w, the EOS, flux5 and the shared pow/sin
 * have closed-form or library checks; the terrain and stretched-eta terms
 * (the large-step PGF and rhs_ph with a varying base geopotential, the
 * terrain surface w) converge at second order to their continuous forms;
 * the w curvature term, w_damping above a vertical CFL of 1, the
 * hypsometric_opt=2 refresh mass and the Coriolis terms have closed forms;
 * the moisture factors have closed forms, zero moisture reduces to the dry
 * step bit for bit, the moist rw_tend and the large-step cqu are pinned at
 * first order in dt, the implicit-w rows by their per-face cq1 scaling and
 * the species' transport by conservation and positivity
 * (test_oracle_moist.py); the factors inside the acoustic loop (cqu/cqv on
 * the u/v pressure gradient, cq1 on the explicit part termA of the implicit
 * w equation), which enter a whole step at second order in dt only, are
 * pinned on the acoustic pieces: with the term's inputs zero the output does
 * not depend on the factor, and otherwise it is affine in each face's own
 * factor (test_oracle_moist_acoustic.py).  What no reference file holds is
 * the reference's own field-level FP64 output beyond the 2-D slabs.
 */
#include "wrfdyn_oracle.h"
#include "wrfdyn_math.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Scalar type of every value the step computes.  The library build uses
 * double.  wrfdyn_gradient.cpp compiles this same file as C++ with ORC_AD
 * defined and `real` a recording scalar, so that its reverse sweep is the
 * derivative of exactly these expressions; that build supplies its own sgn
 * (with a zero derivative, as jnp.sign's) and leaves out the double entry
 * points below. */
#ifndef ORC_AD
typedef double real;
#endif

#define R_D 287.0
/* ORC_LEGACY_V013 rebuilds the oracle with the two dycore constants the
 * reference changed after its v0.13 golden warm-bubble runs (cp 1004 -> 7R/2,
 * acoustic_wrf.py:45-54; the w curvature term added to rw_tend,
 * operational_mode.py:1784-1822), so the oracle can be checked against those
 * reference-held golden scalars. */
#ifdef ORC_LEGACY_V013
#define CP_D 1004.0
#else
#define CP_D (7.0 * R_D / 2.0)
#endif
#define P0_PA 100000.0
#define G 9.81
#define T0 300.0
#define EMDIV 0.01
#define SMDIV 0.1
#define MIN_ACOUSTIC_ALT 0.1
#define MIN_ALT 1.0e-8
#define MIN_PRESSURE 1.0
#define W_ALPHA 0.3
#define W_BETA 1.0

static const double CPOVCV = CP_D / (CP_D - R_D);
static const double CVPM = -(CP_D - R_D) / CP_D;
#ifdef ORC_LEGACY_V013
static const double RERADIUS = 0.0;
#else
static const double RERADIUS = 1.0 / 6370.0e3;
#endif
static const double PI_D = 3.141592653589793;

typedef struct {
    int nx, ny, nz;
} dims;

#define IM(k, j, i) (((size_t)(k) * ny + (size_t)(j)) * nx + (size_t)(i))
#define IU(k, j, i) (((size_t)(k) * ny + (size_t)(j)) * (nx + 1) + (size_t)(i))
#define IV(k, j, i) (((size_t)(k) * (ny + 1) + (size_t)(j)) * nx + (size_t)(i))
#define I2(j, i) ((size_t)(j) * nx + (size_t)(i))
#define I2U(j, i) ((size_t)(j) * (nx + 1) + (size_t)(i))
#define I2V(j, i) ((size_t)(j) * nx + (size_t)(i))

static inline int wrap(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}
static inline int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

#ifndef ORC_AD
/* jnp.sign: -1, +1, and the argument itself for +-0 and NaN. */
static inline real sgn(real x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }
#endif

/* jnp.where(|x|>floor, x, floor) */
static inline real safe_pos(real x, double floor) { return fabs(x) > floor ? x : floor; }

/* The deterministic elementary functions shared with the HIP kernels
 * (include/wrfdyn_math.h), exported for their accuracy tests. */
double orc_pow(double x, double y) { return wd_pow(x, y); }
double orc_sin(double x) { return wd_sin(x); }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Equation of state                                                   */
/* ------------------------------------------------------------------ */

/* _inverse_density_from_theta_pressure, dynamics/acoustic_wrf.py:178-195
 * (qvf = 1.0: state theta is theta_m). */
static real alt_from_theta_p(real theta, real p) {
    real ps = p > MIN_PRESSURE ? p : MIN_PRESSURE;
    return (((R_D / P0_PA) * theta) * 1.0) * wd_pow(ps / P0_PA, CVPM);
}

/* _pressure_from_theta_alt, dynamics/acoustic_wrf.py:198-211. */
static real p_from_theta_alt(real theta, real alt) {
    real as = alt > MIN_ALT ? alt : MIN_ALT;
    real arg = ((R_D * theta) * 1.0) / (P0_PA * as);
    real a2 = arg > 1.0e-12 ? arg : 1.0e-12;
    return P0_PA * wd_pow(a2, CPOVCV);
}

#ifndef ORC_AD
double orc_alt_from_theta_p(double theta, double p) { return alt_from_theta_p(theta, p); }
double orc_p_from_theta_alt(double theta, double alt) { return p_from_theta_alt(theta, alt); }
#endif

/* ------------------------------------------------------------------ */
/* WRF flux operators, dynamics/flux_advection.py:53-88                 */
/* ------------------------------------------------------------------ */

/* The prognostic state as the step computes on it (orc_state's fields). */
typedef struct {
    real *u, *v, *w, *theta;
    real *p_total, *p_pert;
    real *ph_total, *ph_pert;
    real *mu_total, *mu_pert;
    real *mudf;
} rstate;

/* flux5_face_periodic: face value at the left face of cell i. */
static inline real flux5(real q_im3, real q_im2, real q_im1, real q_i, real q_ip1,
                           real q_ip2, real vel) {
    real f6 = ((37.0 * (q_i + q_im1) - 8.0 * (q_ip1 + q_im2)) + (q_ip2 + q_im3)) / 60.0;
    real corr = (((q_ip2 - q_im3) - 5.0 * (q_ip1 - q_im2)) + 10.0 * (q_i - q_im1)) / 60.0;
    return f6 - sgn(vel) * corr;
}

/* flux3 as written in advect_scalar_flux / _vertical_flux_div_3 /
 * _vertical_flux_div_w (flux_advection.py:537-547,1388-1397,1452-1467):
 * flux4 + sign(velz)*corr with the caller choosing velz. */
static inline real flux3(real q_km2, real q_km1, real q_k, real q_kp1, real velz) {
    real f4 = (7.0 * (q_k + q_km1) - (q_kp1 + q_km2)) / 12.0;
    real corr = ((q_kp1 - q_km2) - 3.0 * (q_k - q_km1)) / 12.0;
    return f4 + sgn(velz) * corr;
}

/* ------------------------------------------------------------------ */
/* couple_velocities_periodic, dynamics/flux_advection.py:394-470       */
/* ru, rv are collapsed to (nz, ny, nx); rom is (nz+1, ny, nx).         */
/* ------------------------------------------------------------------ */
static void couple_velocities(const orc_grid *g, const real *u, const real *v,
                              const real *mu, real *ru, real *rv, real *rom,
                              real *divv) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) {
                real muu = 0.5 * (mu[I2(j, i)] + mu[I2(j, wrap(i - 1, nx))]);
                real muv = 0.5 * (mu[I2(j, i)] + mu[I2(wrap(j - 1, ny), i)]);
                real mass_u = g->c1h[k] * muu + g->c2h[k];
                real mass_v = g->c1h[k] * muv + g->c2h[k];
                ru[IM(k, j, i)] = mass_u * u[IU(k, j, i)] / g->msfuy[I2U(j, i)];
                rv[IM(k, j, i)] = mass_v * v[IV(k, j, i)] / g->msfvx[I2V(j, i)];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            real dmdt = 0.0;
            for (int k = 0; k < nz; ++k) {
                real d = (g->dnw[k] * g->msftx[I2(j, i)]) *
                           (rdx * (ru[IM(k, j, wrap(i + 1, nx))] - ru[IM(k, j, i)]) +
                            rdy * (rv[IM(k, wrap(j + 1, ny), i)] - rv[IM(k, j, i)]));
                divv[IM(k, j, i)] = d;
                dmdt += d;
            }
            real cum = 0.0;
            rom[IM(0, j, i)] = 0.0;
            for (int k = 0; k < nz; ++k) {
                real incr = -((g->dnw[k] * g->c1h[k]) * dmdt) - divv[IM(k, j, i)];
                cum = (k == 0) ? incr : cum + incr;
                if (k + 1 < nz) rom[IM(k + 1, j, i)] = cum;
            }
            rom[IM(nz, j, i)] = 0.0;
        }
    }
}

/* Public helper: advect_scalar_flux of `field` with the velocities of (u,v,mu). */
static void advect_scalar_flux(const orc_grid *g, const real *field, const real *ru,
                               const real *rv, const real *rom, real *tend);

#ifndef ORC_AD
int orc_advect_scalar(const orc_grid *g, const double *u, const double *v,
                      const double *mu_total, const double *field, double *tend) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t m = (size_t)nz * ny * nx, w = (size_t)(nz + 1) * ny * nx;
    real *ru = (real *)malloc(m * sizeof(real)), *rv = (real *)malloc(m * sizeof(real));
    real *rom = (real *)malloc(w * sizeof(real)), *dv = (real *)malloc(m * sizeof(real));
    if (!ru || !rv || !rom || !dv) {
        free(ru); free(rv); free(rom); free(dv);
        return -1;
    }
    couple_velocities(g, u, v, mu_total, ru, rv, rom, dv);
    advect_scalar_flux(g, field, ru, rv, rom, tend);
    free(ru); free(rv); free(rom); free(dv);
    return 0;
}
#endif

/* advect_scalar_flux, dynamics/flux_advection.py:482-568 (periodic branch). */
static void advect_scalar_flux(const orc_grid *g, const real *q, const real *ru,
                               const real *rv, const real *rom, real *tend) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) {
#define QX(s) q[IM(k, j, wrap(i + (s), nx))]
#define QY(jj, s) q[IM(k, wrap((jj) + (s), ny), i)]
                int ip = wrap(i + 1, nx);
                real vel0 = ru[IM(k, j, i)], vel1 = ru[IM(k, j, ip)];
                real fq0 = vel0 * flux5(QX(-3), QX(-2), QX(-1), QX(0), QX(1), QX(2), vel0);
                real fq1 = vel1 * flux5(QX(-2), QX(-1), QX(0), QX(1), QX(2), QX(3), vel1);
                real t = ((-g->msftx[I2(j, i)]) * rdx) * (fq1 - fq0);
                int jp = wrap(j + 1, ny);
                real vy0 = rv[IM(k, j, i)], vy1 = rv[IM(k, jp, i)];
                real gy0 = vy0 * flux5(QY(j, -3), QY(j, -2), QY(j, -1), QY(j, 0), QY(j, 1), QY(j, 2), vy0);
                real gy1 = vy1 * flux5(QY(jp, -3), QY(jp, -2), QY(jp, -1), QY(jp, 0), QY(jp, 1), QY(jp, 2), vy1);
                t = t - (g->msftx[I2(j, i)] * rdy) * (gy1 - gy0);
                tend[IM(k, j, i)] = t;
#undef QX
#undef QY
            }
        }
    }
    /* vertical (order 3) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            real vprev = 0.0; /* vflux[0] */
            for (int k = 0; k < nz; ++k) {
                int kf = k + 1; /* upper face */
                real vf;
                if (kf == nz) {
                    vf = 0.0;
                } else if (kf == 1) {
                    vf = rom[IM(1, j, i)] * (g->fnm[1] * q[IM(1, j, i)] + g->fnp[1] * q[IM(0, j, i)]);
                } else if (kf == nz - 1) {
                    vf = rom[IM(nz - 1, j, i)] *
                         (g->fnm[nz - 1] * q[IM(nz - 1, j, i)] + g->fnp[nz - 1] * q[IM(nz - 2, j, i)]);
                } else {
                    real r = rom[IM(kf, j, i)];
                    vf = r * flux3(q[IM(kf - 2, j, i)], q[IM(kf - 1, j, i)], q[IM(kf, j, i)],
                                   q[IM(kf + 1, j, i)], -r);
                }
                tend[IM(k, j, i)] = tend[IM(k, j, i)] - g->rdnw[k] * (vf - vprev);
                vprev = vf;
            }
        }
    }
}

/* _vertical_flux_div_3, dynamics/flux_advection.py:1362-1407 for one column of a
 * mass-level field q (stride sq between levels) with transport romq (stride sr). */
static inline real vflux3_face(const orc_grid *g, int kf, const real *q, size_t sq,
                                 const real *romq, size_t sr) {
    const int nz = g->nz;
    if (kf == 0 || kf == nz) return 0.0;
    if (kf == 1) return romq[1 * sr] * (g->fnm[1] * q[1 * sq] + g->fnp[1] * q[0]);
    if (kf == nz - 1)
        return romq[(size_t)(nz - 1) * sr] *
               (g->fnm[nz - 1] * q[(size_t)(nz - 1) * sq] + g->fnp[nz - 1] * q[(size_t)(nz - 2) * sq]);
    real r = romq[(size_t)kf * sr];
    return r * flux3(q[(size_t)(kf - 2) * sq], q[(size_t)(kf - 1) * sq], q[(size_t)kf * sq],
                     q[(size_t)(kf + 1) * sq], -r);
}

/* advect_u_flux, dynamics/flux_advection.py:1131-1202 (periodic branch).
 * u_t is (nz, ny, nx+1) with face nx restored from face 0. */
static void advect_u_flux(const orc_grid *g, const real *u, const real *ru, const real *rv,
                          const real *rom, real *ut) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) {
#define UX(s) u[IU(k, j, wrap(i + (s), nx))]
#define UY(jj, s) u[IU(k, wrap((jj) + (s), ny), i)]
                int ip = wrap(i + 1, nx), im = wrap(i - 1, nx);
                real velx0 = 0.5 * (ru[IM(k, j, i)] + ru[IM(k, j, im)]);
                real velx1 = 0.5 * (ru[IM(k, j, ip)] + ru[IM(k, j, i)]);
                real fq0 = velx0 * flux5(UX(-3), UX(-2), UX(-1), UX(0), UX(1), UX(2), velx0);
                real fq1 = velx1 * flux5(UX(-2), UX(-1), UX(0), UX(1), UX(2), UX(3), velx1);
                real msf = g->msfux[I2U(j, i)];
                real t = ((-msf) * rdx) * (fq1 - fq0);
                if (ny > 1) {
                    int jp = wrap(j + 1, ny);
                    real vy0 = 0.5 * (rv[IM(k, j, i)] + rv[IM(k, j, im)]);
                    real vy1 = 0.5 * (rv[IM(k, jp, i)] + rv[IM(k, jp, im)]);
                    real g0 = vy0 * flux5(UY(j, -3), UY(j, -2), UY(j, -1), UY(j, 0), UY(j, 1), UY(j, 2), vy0);
                    real g1 = vy1 * flux5(UY(jp, -3), UY(jp, -2), UY(jp, -1), UY(jp, 0), UY(jp, 1), UY(jp, 2), vy1);
                    t = t - (msf * rdy) * (g1 - g0);
                }
                ut[IU(k, j, i)] = t;
#undef UX
#undef UY
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        real *romq = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        real *col = (real *)malloc((size_t)nz * sizeof(real));
        for (int i = 0; i < nx; ++i) {
            int im = wrap(i - 1, nx);
            for (int k = 0; k <= nz; ++k) romq[k] = 0.5 * (rom[IM(k, j, i)] + rom[IM(k, j, im)]);
            for (int k = 0; k < nz; ++k) col[k] = u[IU(k, j, i)];
            real vprev = vflux3_face(g, 0, col, 1, romq, 1);
            for (int k = 0; k < nz; ++k) {
                real vf = vflux3_face(g, k + 1, col, 1, romq, 1);
                real vd = (-g->rdnw[k]) * (vf - vprev);
                ut[IU(k, j, i)] = ut[IU(k, j, i)] + vd;
                vprev = vf;
            }
        }
        free(romq);
        free(col);
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) ut[IU(k, j, nx)] = ut[IU(k, j, 0)];
}

/* advect_v_flux, dynamics/flux_advection.py:1205-1274 (periodic branch).
 * v_t is (nz, ny+1, nx) with row ny restored from row 0. */
static void advect_v_flux(const orc_grid *g, const real *v, const real *ru, const real *rv,
                          const real *rom, real *vt) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        int jm = wrap(j - 1, ny), jp = wrap(j + 1, ny), jpm = j; /* row jp's south neighbour is j */
        (void)jpm;
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) {
#define VX(s) v[IV(k, j, wrap(i + (s), nx))]
#define VY(jj, s) v[IV(k, wrap((jj) + (s), ny), i)]
                int ip = wrap(i + 1, nx);
                real vx0 = 0.5 * (ru[IM(k, j, i)] + ru[IM(k, jm, i)]);
                real vx1 = 0.5 * (ru[IM(k, j, ip)] + ru[IM(k, jm, ip)]);
                real fq0 = vx0 * flux5(VX(-3), VX(-2), VX(-1), VX(0), VX(1), VX(2), vx0);
                real fq1 = vx1 * flux5(VX(-2), VX(-1), VX(0), VX(1), VX(2), VX(3), vx1);
                real msfvy = g->msfvy[I2V(j, i)];
                real t = ((-msfvy) * rdx) * (fq1 - fq0);
                real vy0 = 0.5 * (rv[IM(k, j, i)] + rv[IM(k, jm, i)]);
                real vy1 = 0.5 * (rv[IM(k, jp, i)] + rv[IM(k, j, i)]);
                real g0 = vy0 * flux5(VY(j, -3), VY(j, -2), VY(j, -1), VY(j, 0), VY(j, 1), VY(j, 2), vy0);
                real g1 = vy1 * flux5(VY(jp, -3), VY(jp, -2), VY(jp, -1), VY(jp, 0), VY(jp, 1), VY(jp, 2), vy1);
                t = t - (msfvy * rdy) * (g1 - g0);
                vt[IV(k, j, i)] = t;
#undef VX
#undef VY
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        int jm = wrap(j - 1, ny);
        real *romq = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        real *col = (real *)malloc((size_t)nz * sizeof(real));
        for (int i = 0; i < nx; ++i) {
            for (int k = 0; k <= nz; ++k) romq[k] = 0.5 * (rom[IM(k, j, i)] + rom[IM(k, jm, i)]);
            for (int k = 0; k < nz; ++k) col[k] = v[IV(k, j, i)];
            real ratio = g->msfvy[I2V(j, i)] / g->msfvx[I2V(j, i)];
            real vprev = vflux3_face(g, 0, col, 1, romq, 1);
            for (int k = 0; k < nz; ++k) {
                real vf = vflux3_face(g, k + 1, col, 1, romq, 1);
                real vd = (-g->rdnw[k]) * (vf - vprev);
                vt[IV(k, j, i)] = vt[IV(k, j, i)] + ratio * vd;
                vprev = vf;
            }
        }
        free(romq);
        free(col);
    }
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) vt[IV(k, ny, i)] = vt[IV(k, 0, i)];
}

/* _mass_to_full_levels, dynamics/flux_advection.py:1332-1348 */
static inline real mass_to_full(const orc_grid *g, const real *r, int kf, int j, int i) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (kf == 0) return r[IM(0, j, i)];
    if (kf == nz) return r[IM(nz - 1, j, i)];
    return g->fnm[kf] * r[IM(kf, j, i)] + g->fnp[kf] * r[IM(kf - 1, j, i)];
}

/* advect_w_flux + _vertical_flux_div_w (top_lid=True),
 * dynamics/flux_advection.py:1277-1329, :1410-1482. */
static void advect_w_flux(const orc_grid *g, const real *w, const real *ru, const real *rv,
                          const real *rom, real *wt) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int k = 0; k <= nz; ++k) {
            for (int i = 0; i < nx; ++i) {
#define WX(s) w[IM(k, j, wrap(i + (s), nx))]
#define WY(jj, s) w[IM(k, wrap((jj) + (s), ny), i)]
                int ip = wrap(i + 1, nx);
                real vx0 = mass_to_full(g, ru, k, j, i), vx1 = mass_to_full(g, ru, k, j, ip);
                real fq0 = vx0 * flux5(WX(-3), WX(-2), WX(-1), WX(0), WX(1), WX(2), vx0);
                real fq1 = vx1 * flux5(WX(-2), WX(-1), WX(0), WX(1), WX(2), WX(3), vx1);
                real msf = g->msftx[I2(j, i)];
                real t = ((-msf) * rdx) * (fq1 - fq0);
                if (ny > 1) {
                    int jp = wrap(j + 1, ny);
                    real vy0 = mass_to_full(g, rv, k, j, i), vy1 = mass_to_full(g, rv, k, jp, i);
                    real g0 = vy0 * flux5(WY(j, -3), WY(j, -2), WY(j, -1), WY(j, 0), WY(j, 1), WY(j, 2), vy0);
                    real g1 = vy1 * flux5(WY(jp, -3), WY(jp, -2), WY(jp, -1), WY(jp, 0), WY(jp, 1), WY(jp, 2), vy1);
                    t = t - (msf * rdy) * (g1 - g0);
                }
                wt[IM(k, j, i)] = t;
#undef WX
#undef WY
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        real *vflux = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        for (int i = 0; i < nx; ++i) {
#define VF(k) (0.5 * (rom[IM((k), j, i)] + rom[IM((k) - 1, j, i)]))
#define WW(k) w[IM((k), j, i)]
            for (int k = 0; k <= nz; ++k) vflux[k] = 0.0;
            for (int k = 3; k < nz; ++k) {
                real vel = VF(k);
                vflux[k] = vel * flux3(WW(k - 2), WW(k - 1), WW(k), WW(k + 1), -vel);
            }
            vflux[1] = (VF(1) * 0.5) * (WW(1) + WW(0));
            {
                real vel = VF(2);
                vflux[2] = vel * flux3(WW(0), WW(1), WW(2), WW(3), -vel);
                vel = VF(nz - 1);
                vflux[nz - 1] = vel * flux3(WW(nz - 3), WW(nz - 2), WW(nz - 1), WW(nz), -vel);
            }
            /* top_lid: vflux[nz] = 0, no pickup */
            for (int k = 1; k < nz; ++k) {
                real vd = (-g->rdn[k]) * (vflux[k + 1] - vflux[k]);
                wt[IM(k, j, i)] = wt[IM(k, j, i)] + vd;
            }
            wt[IM(0, j, i)] = wt[IM(0, j, i)] + 0.0;
            wt[IM(nz, j, i)] = wt[IM(nz, j, i)] + 0.0;
#undef VF
#undef WW
        }
        free(vflux);
    }
}

/* ------------------------------------------------------------------ */
/* Large-step horizontal PGF                                            */
/* core/rk_addtend_dry.py:127-229 (_absolute_diagnostics, hyps opt 1)   */
/* core/rk_addtend_dry.py:232-361 (large_step_horizontal_pgf)          */
/* ------------------------------------------------------------------ */
/* hypsometric_opt 2 (WRF calc_p_rho_phi log form, rk_addtend_dry.py:196-222,
 * acoustic_wrf.py:394-427): the specific volume of layer k from its
 * geopotential depth dph and the column mass: dph / phm / ln(pfd / pfu), with
 * the face and half-level dry pressures c3 mass + c4 + p_top. */
static inline real log_alpha(const orc_grid *g, int k, real dph, real mass) {
    const real pfu = g->c3f[k + 1] * mass + g->c4f[k + 1] + g->p_top;
    const real pfd = g->c3f[k] * mass + g->c4f[k] + g->p_top;
    const real phm = g->c3h[k] * mass + g->c4h[k] + g->p_top;
    return dph / phm / wd_log(pfd / pfu);
}

/* cqu, cqv: the moisture factors on u and v faces (ru_tend -= cqu dpx,
 * rk_addtend_dry.py:278, :359-360); NULL for the dry state's 1. */
static void large_step_pgf(const orc_grid *g, int hypsometric_opt, const rstate *S, const real *alt,
                           real *al, real *php, real *ut, real *vt, real *dpn_col, const real *cqu,
                           const real *cqv) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    real dn_top = g->dn[nz - 1];
    real dn_safe = fabs(dn_top) > 1.0e-30 ? dn_top : 1.0;
    real cfn = (0.5 * g->dnw[nz - 1] + g->dn[nz - 1]) / dn_safe;
    real cfn1 = -0.5 * g->dnw[nz - 1] / dn_safe;
    (void)dpn_col;

    /* _absolute_diagnostics: al, php */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            real mu_pert = S->mu_pert[I2(j, i)];
            real mu_total = S->mu_total[I2(j, i)];
            real mub = mu_total - mu_pert;
            real muts = mu_total + mu_pert;
            for (int k = 0; k < nz; ++k) {
                real phb_k = S->ph_total[IM(k, j, i)] - S->ph_pert[IM(k, j, i)];
                real phb_k1 = S->ph_total[IM(k + 1, j, i)] - S->ph_pert[IM(k + 1, j, i)];
                real mass_base = g->c1h[k] * mub + g->c2h[k];
                real safe_base = safe_pos(mass_base, 1.0e-12);
                real alb = ((-g->rdnw[k]) * (phb_k1 - phb_k)) / safe_base;
                if (hypsometric_opt == 2) {
                    /* muts here is the state's mu_total (rk_addtend_dry.py:207) */
                    real dph_tot = (phb_k1 + S->ph_pert[IM(k + 1, j, i)]) - (phb_k + S->ph_pert[IM(k, j, i)]);
                    al[IM(k, j, i)] = log_alpha(g, k, dph_tot, mu_total) - log_alpha(g, k, phb_k1 - phb_k, mub);
                } else {
                    real mass_t = g->c1h[k] * muts + g->c2h[k];
                    real safe_t = safe_pos(mass_t, 1.0e-12);
                    real mu_term = g->c1h[k] * mu_pert;
                    al[IM(k, j, i)] =
                        -(alb * mu_term + g->rdnw[k] * (S->ph_pert[IM(k + 1, j, i)] - S->ph_pert[IM(k, j, i)])) /
                        safe_t;
                }
                real pht_k = phb_k + S->ph_pert[IM(k, j, i)];
                real pht_k1 = phb_k1 + S->ph_pert[IM(k + 1, j, i)];
                php[IM(k, j, i)] = 0.5 * (pht_k + pht_k1);
            }
        }
    }

    /* x PGF on u faces (edge-padded face pairs) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        real *dpn = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        for (int i = 0; i <= nx; ++i) {
            int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
            real msf_u = g->msfux[I2U(j, i)] / g->msfuy[I2U(j, i)];
            real mub_l = S->mu_total[I2(j, il)] - S->mu_pert[I2(j, il)];
            real mub_r = S->mu_total[I2(j, ir)] - S->mu_pert[I2(j, ir)];
            real mt_l = mub_l + S->mu_pert[I2(j, il)];
            real mt_r = mub_r + S->mu_pert[I2(j, ir)];
            real muu = 0.5 * (mt_l + mt_r);
            real mu_l = S->mu_pert[I2(j, il)], mu_r = S->mu_pert[I2(j, ir)];
            /* dpn faces from pair sums of p' */
            {
                real ps0 = S->p_pert[IM(0, j, il)] + S->p_pert[IM(0, j, ir)];
                real ps1 = S->p_pert[IM(1, j, il)] + S->p_pert[IM(1, j, ir)];
                real ps2 = S->p_pert[IM(2, j, il)] + S->p_pert[IM(2, j, ir)];
                dpn[0] = 0.5 * ((g->cf1 * ps0 + g->cf2 * ps1) + g->cf3 * ps2);
                for (int k = 1; k < nz; ++k) {
                    real psk = S->p_pert[IM(k, j, il)] + S->p_pert[IM(k, j, ir)];
                    real pskm = S->p_pert[IM(k - 1, j, il)] + S->p_pert[IM(k - 1, j, ir)];
                    dpn[k] = 0.5 * (g->fnm[k] * psk + g->fnp[k] * pskm);
                }
                real pst = S->p_pert[IM(nz - 1, j, il)] + S->p_pert[IM(nz - 1, j, ir)];
                real pstm = S->p_pert[IM(nz - 2, j, il)] + S->p_pert[IM(nz - 2, j, ir)];
                dpn[nz] = 0.5 * (cfn * pst + cfn1 * pstm);
            }
            for (int k = 0; k < nz; ++k) {
                real mass_u = g->c1h[k] * muu + g->c2h[k];
                real ph_term = (S->ph_pert[IM(k + 1, j, ir)] - S->ph_pert[IM(k + 1, j, il)]) +
                                 (S->ph_pert[IM(k, j, ir)] - S->ph_pert[IM(k, j, il)]);
                real p_term = (alt[IM(k, j, il)] + alt[IM(k, j, ir)]) *
                                (S->p_pert[IM(k, j, ir)] - S->p_pert[IM(k, j, il)]);
                real pb_l = S->p_total[IM(k, j, il)] - S->p_pert[IM(k, j, il)];
                real pb_r = S->p_total[IM(k, j, ir)] - S->p_pert[IM(k, j, ir)];
                real pb_term = (al[IM(k, j, il)] + al[IM(k, j, ir)]) * (pb_r - pb_l);
                real dpx = (((msf_u * 0.5) * rdx) * mass_u) * ((ph_term + p_term) + pb_term);
                real bracket = g->rdnw[k] * (dpn[k + 1] - dpn[k]) - 0.5 * (g->c1h[k] * (mu_l + mu_r));
                dpx = dpx + ((msf_u * rdx) * (php[IM(k, j, ir)] - php[IM(k, j, il)])) * bracket;
                ut[IU(k, j, i)] = ut[IU(k, j, i)] + (-(cqu ? cqu[IU(k, j, i)] : (real)1.0)) * dpx;
            }
        }
        free(dpn);
    }

    /* y PGF on v faces */
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny; ++j) {
        real *dpn = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
        for (int i = 0; i < nx; ++i) {
            real msf_v = g->msfvy[I2V(j, i)] / g->msfvx[I2V(j, i)];
            real mub_s = S->mu_total[I2(js, i)] - S->mu_pert[I2(js, i)];
            real mub_n = S->mu_total[I2(jn, i)] - S->mu_pert[I2(jn, i)];
            real mt_s = mub_s + S->mu_pert[I2(js, i)];
            real mt_n = mub_n + S->mu_pert[I2(jn, i)];
            real muv = 0.5 * (mt_s + mt_n);
            real mu_s = S->mu_pert[I2(js, i)], mu_n = S->mu_pert[I2(jn, i)];
            {
                real ps0 = S->p_pert[IM(0, js, i)] + S->p_pert[IM(0, jn, i)];
                real ps1 = S->p_pert[IM(1, js, i)] + S->p_pert[IM(1, jn, i)];
                real ps2 = S->p_pert[IM(2, js, i)] + S->p_pert[IM(2, jn, i)];
                dpn[0] = 0.5 * ((g->cf1 * ps0 + g->cf2 * ps1) + g->cf3 * ps2);
                for (int k = 1; k < nz; ++k) {
                    real psk = S->p_pert[IM(k, js, i)] + S->p_pert[IM(k, jn, i)];
                    real pskm = S->p_pert[IM(k - 1, js, i)] + S->p_pert[IM(k - 1, jn, i)];
                    dpn[k] = 0.5 * (g->fnm[k] * psk + g->fnp[k] * pskm);
                }
                real pst = S->p_pert[IM(nz - 1, js, i)] + S->p_pert[IM(nz - 1, jn, i)];
                real pstm = S->p_pert[IM(nz - 2, js, i)] + S->p_pert[IM(nz - 2, jn, i)];
                dpn[nz] = 0.5 * (cfn * pst + cfn1 * pstm);
            }
            for (int k = 0; k < nz; ++k) {
                real mass_v = g->c1h[k] * muv + g->c2h[k];
                real ph_term = (S->ph_pert[IM(k + 1, jn, i)] - S->ph_pert[IM(k + 1, js, i)]) +
                                 (S->ph_pert[IM(k, jn, i)] - S->ph_pert[IM(k, js, i)]);
                real p_term = (alt[IM(k, js, i)] + alt[IM(k, jn, i)]) *
                                (S->p_pert[IM(k, jn, i)] - S->p_pert[IM(k, js, i)]);
                real pb_s = S->p_total[IM(k, js, i)] - S->p_pert[IM(k, js, i)];
                real pb_n = S->p_total[IM(k, jn, i)] - S->p_pert[IM(k, jn, i)];
                real pb_term = (al[IM(k, js, i)] + al[IM(k, jn, i)]) * (pb_n - pb_s);
                real dpy = (((msf_v * 0.5) * rdy) * mass_v) * ((ph_term + p_term) + pb_term);
                real bracket = g->rdnw[k] * (dpn[k + 1] - dpn[k]) - 0.5 * (g->c1h[k] * (mu_s + mu_n));
                dpy = dpy + ((msf_v * rdy) * (php[IM(k, jn, i)] - php[IM(k, js, i)])) * bracket;
                vt[IV(k, j, i)] = vt[IV(k, j, i)] + (-(cqv ? cqv[IV(k, j, i)] : (real)1.0)) * dpy;
            }
        }
        free(dpn);
    }
}

/* ------------------------------------------------------------------ */
/* large_step_coriolis, dynamics/core/rk_addtend_dry.py:364-483, added  */
/* to the coupled u/v tendencies right after the large-step PGF        */
/* (operational_mode.py:3052-3058) with specified = run_boundary =     */
/* false.  The coupled momenta are rebuilt here from the stage state   */
/* with edge-padded face masses of mu_total (:415-419), not taken from */
/* couple_velocities (periodic face masses):                           */
/*   ru_cor = (msfux/msfuy) f_u rv_quad - e_u cosa_u rw_quad_x         */
/*   rv_cor = -(msfvy/msfvx) f_v ru_quad + (msfvy/msfvx) e_v sina_v     */
/*            rw_quad_y                                                 */
/* with the f/e/sina/cosa face averages of the two mass points and the */
/* off-axis momenta averaged over the four surrounding faces, x and y  */
/* neighbours clamped at the array edges (:436-470).                   */
/* ------------------------------------------------------------------ */
static inline real cor_ru(const orc_grid *g, const rstate *S, int k, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    (void)ny;
    const int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
    real muu = 0.5 * (S->mu_total[I2(j, il)] + S->mu_total[I2(j, ir)]);
    return (S->u[IU(k, j, i)] * (g->c1h[k] * muu + g->c2h[k])) / g->msfuy[I2U(j, i)];
}
static inline real cor_rv(const orc_grid *g, const rstate *S, int k, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    const int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
    real muv = 0.5 * (S->mu_total[I2(js, i)] + S->mu_total[I2(jn, i)]);
    return (S->v[IV(k, j, i)] * (g->c1h[k] * muv + g->c2h[k])) / g->msfvx[I2V(j, i)];
}
static inline real cor_rw(const orc_grid *g, const rstate *S, int k, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    (void)ny;
    return (S->w[IM(k, j, i)] * (g->c1f[k] * S->mu_total[I2(j, i)] + g->c2f[k])) / g->msfty[I2(j, i)];
}

static void large_step_coriolis(const orc_grid *g, const rstate *S, real *ut, real *vt) {
    if (!g->f) return;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
                real msf_u = g->msfux[I2U(j, i)] / g->msfuy[I2U(j, i)];
                real f_u = 0.5 * (g->f[I2(j, il)] + g->f[I2(j, ir)]);
                real e_u = 0.5 * (g->e[I2(j, il)] + g->e[I2(j, ir)]);
                real cosa_u = 0.5 * (g->cosa[I2(j, il)] + g->cosa[I2(j, ir)]);
                real rv_quad = 0.25 * (((cor_rv(g, S, k, j, il) + cor_rv(g, S, k, j, ir)) +
                                        cor_rv(g, S, k, j + 1, il)) +
                                       cor_rv(g, S, k, j + 1, ir));
                real ru_cor = (msf_u * f_u) * rv_quad;
                real rw_quad = 0.25 * (((cor_rw(g, S, k, j, il) + cor_rw(g, S, k + 1, j, il)) +
                                        cor_rw(g, S, k, j, ir)) +
                                       cor_rw(g, S, k + 1, j, ir));
                ru_cor = ru_cor - (e_u * cosa_u) * rw_quad;
                ut[IU(k, j, i)] = ut[IU(k, j, i)] + ru_cor;
            }
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
                real msf_v = g->msfvy[I2V(j, i)] / g->msfvx[I2V(j, i)];
                real f_v = 0.5 * (g->f[I2(js, i)] + g->f[I2(jn, i)]);
                real e_v = 0.5 * (g->e[I2(js, i)] + g->e[I2(jn, i)]);
                real sina_v = 0.5 * (g->sina[I2(js, i)] + g->sina[I2(jn, i)]);
                real ru_quad = 0.25 * (((cor_ru(g, S, k, js, i) + cor_ru(g, S, k, js, i + 1)) +
                                        cor_ru(g, S, k, jn, i)) +
                                       cor_ru(g, S, k, jn, i + 1));
                real rv_cor = ((-msf_v) * f_v) * ru_quad;
                real rw_quad = 0.25 * (((cor_rw(g, S, k, js, i) + cor_rw(g, S, k + 1, js, i)) +
                                        cor_rw(g, S, k, jn, i)) +
                                       cor_rw(g, S, k + 1, jn, i));
                rv_cor = rv_cor + ((msf_v * e_v) * sina_v) * rw_quad;
                vt[IV(k, j, i)] = vt[IV(k, j, i)] + rv_cor;
            }
    }
}

/* The cosine-Coriolis term of the stage rw_tend at interior face k
 * (operational_mode.py:1818-1821): e (cosa ru_f - (msftx/msfty) sina rv_f)
 * with the face-averaged coupled momenta ru_f, rv_f the curvature term uses. */
static inline real cor_w(const orc_grid *g, size_t c, real mxy, real ru_f, real rv_f) {
    if (!g->f) return 0.0;
    return g->e[c] * (g->cosa[c] * ru_f - (mxy * g->sina[c]) * rv_f);
}

/* ------------------------------------------------------------------ */
/* calc_coef_w_wrf_coefficients, dynamics/acoustic_wrf.py:783-852       */
/* cqw is the post-pg_buoy_w field: 0 at faces 0 and nz, 1 inside       */
/* (core/advance_w.py:266-290 with qtot = 0).                           */
/* ------------------------------------------------------------------ */
static inline real cqw_at(int kf, int nz) { return (kf >= 1 && kf < nz) ? 1.0 : 0.0; }
/* cq1 = 1/(1 + cqw) of pg_buoy_w_moist (core/advance_w.py:266-290) from the
 * raw calc_cq face value; 1 for the dry state's cqw = 0. */
static inline real cq1_of(real raw) { return 1.0 / (1.0 + raw); }

static void calc_coef_w_column(const orc_grid *g, real mut, const real *c2a, size_t sc,
                               double dts, double epssm, int top_lid, real *a, real *alpha,
                               real *gamma, size_t sw, const real *cqwr, size_t scq) {
    const int nz = g->nz;
    /* the post-pg_buoy_w cqw: cq1 inside, 0 at faces 0 and nz; cqwr (the raw
     * calc_cq field of the column, stride scq) NULL for the dry state */
#define CQW(k) (((k) >= 1 && (k) < nz) ? (cqwr ? cq1_of(cqwr[(size_t)(k) * scq]) : cqw_at((k), nz)) : (real)0.0)
    real cof = (0.5 * dts * G * (1.0 + epssm));
    cof = cof * cof;
    real lid_flag = top_lid ? 0.0 : 1.0;
#define MH(k) (g->c1h[(k)] * mut + g->c2h[(k)])
#define MF(k) (g->c1f[(k)] * mut + g->c2f[(k)])
#define C2A(k) c2a[(size_t)(k) * sc]
    for (int k = 0; k <= nz; ++k) {
        a[(size_t)k * sw] = 0.0;
        alpha[(size_t)k * sw] = 1.0;
        gamma[(size_t)k * sw] = 0.0;
    }
    real top_denom_a = MH(nz - 1) * MF(nz - 1);
    real top_denom_b = MH(nz - 1) * MF(nz);
    real r2 = g->rdnw[nz - 1] * g->rdnw[nz - 1];
    a[1 * sw] = 0.0;
    a[(size_t)nz * sw] = ((((-2.0 * cof) * r2) * C2A(nz - 1)) * lid_flag) / top_denom_a;
    gamma[0] = 0.0;
    for (int kk = 2; kk < nz; ++kk) {
        int k = kk - 1;
        real denom = MH(k) * MF(k);
        a[(size_t)kk * sw] =
            (((((-CQW(kk)) * cof) * g->rdn[kk]) * g->rdnw[kk - 1]) * C2A(kk - 1)) / denom;
    }
    for (int k = 1; k < nz; ++k) {
        real du = MH(k) * MF(k);
        real dl = MH(k - 1) * MF(k);
        real dc = MH(k) * MF(k + 1);
        real b = 1.0 + ((CQW(k) * cof) * g->rdn[k]) *
                             ((g->rdnw[k] * C2A(k)) / du + (g->rdnw[k - 1] * C2A(k - 1)) / dl);
        real c = (((((-CQW(k)) * cof) * g->rdn[k]) * g->rdnw[k]) * C2A(k)) / dc;
        real al = 1.0 / (b - a[(size_t)k * sw] * gamma[(size_t)(k - 1) * sw]);
        alpha[(size_t)k * sw] = al;
        gamma[(size_t)k * sw] = c * al;
    }
    real b_top = 1.0 + (((2.0 * cof) * r2) * C2A(nz - 1)) / top_denom_b;
    alpha[(size_t)nz * sw] = 1.0 / (b_top - a[(size_t)nz * sw] * gamma[(size_t)(nz - 1) * sw]);
    gamma[(size_t)nz * sw] = 0.0;
#undef MH
#undef MF
#undef C2A
#undef CQW
}

#ifndef ORC_AD
int orc_calc_coef_w_cq(const orc_grid *g, const double *mut, const double *c2a, double dts, double epssm,
                       int top_lid, const double *cqw_raw, double *a, double *alpha, double *gamma) {
    const int nx = g->nx, ny = g->ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            calc_coef_w_column(g, mut[I2(j, i)], c2a + I2(j, i), (size_t)ny * nx, dts, epssm, top_lid,
                               a + I2(j, i), alpha + I2(j, i), gamma + I2(j, i), (size_t)ny * nx,
                               cqw_raw ? cqw_raw + I2(j, i) : NULL, (size_t)ny * nx);
    return 0;
}

int orc_calc_coef_w(const orc_grid *g, const double *mut, const double *c2a, double dts,
                    double epssm, int top_lid, double *a, double *alpha, double *gamma) {
    return orc_calc_coef_w_cq(g, mut, c2a, dts, epssm, top_lid, NULL, a, alpha, gamma);
}

/* Thomas sweeps, core/advance_w.py:558-592 and dynamics/tridiag_solve.py:15-57. */
void orc_thomas(int nzp1, const double *a, const double *alpha, const double *gamma,
                const double *rhs, double *out) {
    int nz = nzp1 - 1;
    out[0] = rhs[0];
    for (int k = 1; k <= nz; ++k) out[k] = (rhs[k] - a[k] * out[k - 1]) * alpha[k];
    for (int k = nz - 1; k >= 1; --k) out[k] = out[k] - gamma[k] * out[k + 1];
}
#endif

/* ------------------------------------------------------------------ */
/* 6th-order numerical filter, dynamics/explicit_diffusion.py:159-223   */
/* (_dflux6, _sixth_axis, sixth_order_diffusion_tendency), applied as   */
/* operational_mode.py:2870-2876 does when diff_6th_opt != 0: periodic */
/* along each horizontal axis over the field's own extent there, with  */
/* the monotonic limiter (the reference passes monotonic=True for any  */
/* nonzero diff_6th_opt).                                              */
/* ------------------------------------------------------------------ */

/* coef (dflux_p1 - dflux_p0) at index i of a periodic line of n values
 * spaced `s` apart starting at f. */
static inline real sixth_axis(const real *f, int i, int n, size_t s, double coef) {
    const real fm3 = f[(size_t)wrap(i - 3, n) * s], fm2 = f[(size_t)wrap(i - 2, n) * s];
    const real fm1 = f[(size_t)wrap(i - 1, n) * s], f0 = f[(size_t)i * s];
    const real fp1 = f[(size_t)wrap(i + 1, n) * s], fp2 = f[(size_t)wrap(i + 2, n) * s];
    const real fp3 = f[(size_t)wrap(i + 3, n) * s];
    real d0 = (10.0 * (f0 - fm1) - 5.0 * (fp1 - fm2)) + (fp2 - fm3);
    real d1 = (10.0 * (fp1 - f0) - 5.0 * (fp2 - fm1)) + (fp3 - fm2);
    const real g0 = f0 - fm1, g1 = fp1 - f0;
    if (d0 * g0 <= 0.0) d0 = 0.0;
    if (d1 * g1 <= 0.0) d1 = 0.0;
    return coef * (d1 - d0);
}

/* The uncoupled tendency of f (nlev, nrow, ncol): x part, plus the y part
 * when there is more than one row. */
static inline real sixth_at(const real *f, int nrow, int ncol, int k, int j, int i, double coef) {
    const size_t row = ((size_t)k * nrow + (size_t)j) * ncol;
    real t = sixth_axis(f + row, i, ncol, 1, coef);
    if (nrow > 1) t = t + sixth_axis(f + (size_t)k * nrow * ncol + (size_t)i, j, nrow, (size_t)ncol, coef);
    return t;
}

/* tend += (c1[k] mu[j, i] + c2[k]) * sixth(f): the coupled tendency. */
static void sixth_filter(const real *f, int nlev, int nrow, int ncol, double coef, const double *c1,
                         const double *c2, const real *mu, real *tend) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nlev; ++k)
        for (int j = 0; j < nrow; ++j)
            for (int i = 0; i < ncol; ++i) {
                const size_t q = ((size_t)k * nrow + (size_t)j) * ncol + (size_t)i;
                const real mass = c1[k] * mu[(size_t)j * ncol + (size_t)i] + c2[k];
                tend[q] = tend[q] + mass * sixth_at(f, nrow, ncol, k, j, i, coef);
            }
}

#ifndef ORC_AD
int orc_sixth_tendency(int nlev, int nrow, int ncol, const double *f, double coef, double *out) {
    if (nlev < 1 || nrow < 1 || ncol < 1 || !f || !out) return -1;
    for (int k = 0; k < nlev; ++k)
        for (int j = 0; j < nrow; ++j)
            for (int i = 0; i < ncol; ++i)
                out[((size_t)k * nrow + (size_t)j) * ncol + (size_t)i] = sixth_at(f, nrow, ncol, k, j, i, coef);
    return 0;
}
#endif

/* ------------------------------------------------------------------ */
/* Workspace                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    /* origin (rk1 reference) copies */
    real *u0, *v0, *w0, *th0, *php0_unused, *ph0, *mu0;
    /* stage velocities */
    real *ru, *rv, *rom, *divv;
    /* tendencies */
    real *ut, *vt, *wt, *tht;
    /* stage diagnostics */
    real *alt_raw, *al_ls, *php_ls;
    /* prep */
    real *mub, *mut, *muts0, *mu_work0, *muu, *muus, *muv, *muvs;
    real *u_work, *v_work, *w_work, *th_work, *ph_work;
    real *c2a, *alt, *pb, *phb, *php;
    real *p, *al, *pm1;
    real *rw_tend, *ph_tend, *ca, *calpha, *cgamma;
    /* acoustic carried */
    real *mu, *muts, *muave, *ww, *t2ave;
    /* substep scratch */
    real *dvdxi, *mudf_new, *ww_new, *wsolve, *rhs;
    /* finish */
    real *tmp_ph_total;
    /* step-origin mu_total for the limited theta advection; outside the
     * array table (NULL in the piece API, which refuses scalar_adv_opt) */
    real *mut0;
    /* the moisture species and their work arrays (NULL: the dry state; always
     * NULL in the piece API) */
    struct moist_s *moist;
} ws_t;

/* The six moisture species qv, qc, qr, qi, qs, qg (operational_mode.py:3100)
 * and the work arrays of their coupling and transport. */
#define N_MOIST 6
typedef struct moist_s {
    real *q[N_MOIST];   /* the state's species, advanced in place */
    real *q0[N_MOIST];  /* step-origin copies (WRF moist_old) */
    real *qt[N_MOIST];  /* the stage's coupled tendencies d(mu q)/dt */
    real *qtot;         /* total moisture on mass levels */
    real *cqu, *cqv;    /* moist factors on u and v faces */
    real *cqwr;         /* raw calc_cq cqw on w faces (0 at faces 0 and nz) */
} moist_t;

/* moisture_coupling_factors (dynamics/acoustic_wrf.py:303-319) and
 * moist_cqw_calc_face (core/advance_w.py:195-219) of the stage-entry
 * moisture: qtot = qv + qc + qr + qi + qs + qg on mass levels;
 * cqu = 1/(1 + (qtot_left + qtot_right)/2) on u faces and cqv likewise on v
 * faces, the neighbours clamped at the array edges; the raw cqw =
 * (qtot(k) + qtot(k-1))/2 on the interior w faces. */
static void moist_factors(const orc_grid *g, moist_t *m) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t Mn = (size_t)nz * ny * nx;
    for (size_t q = 0; q < Mn; ++q)
        m->qtot[q] = ((((m->q[0][q] + m->q[1][q]) + m->q[2][q]) + m->q[3][q]) + m->q[4][q]) + m->q[5][q];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
                m->cqu[IU(k, j, i)] = 1.0 / (1.0 + 0.5 * (m->qtot[IM(k, j, il)] + m->qtot[IM(k, j, ir)]));
            }
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
                m->cqv[IV(k, j, i)] = 1.0 / (1.0 + 0.5 * (m->qtot[IM(k, js, i)] + m->qtot[IM(k, jn, i)]));
            }
    }
    for (int f = 0; f <= nz; ++f)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                m->cqwr[IM(f, j, i)] =
                    (f >= 1 && f < nz) ? 0.5 * (m->qtot[IM(f, j, i)] + m->qtot[IM(f - 1, j, i)]) : (real)0.0;
}

/* Terrain surface w of advance_w (core/advance_w.py:505-556): the kinematic
 * condition from u and v extrapolated to the surface (cf1..cf3 over mass
 * levels 0..2) and the edge-clamped terrain height differences, the height
 * being the surface base geopotential phb0 / g.  phb0 is the level-0 plane
 * (ny*nx). */
static real surface_w(const orc_grid *g, const real *phb0, const real *u, const real *v, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    const real msftx = g->msftx[I2(j, i)], msfty = g->msfty[I2(j, i)];
    int jn = clampi(j + 1, 0, ny - 1), js = clampi(j - 1, 0, ny - 1);
    int ie = clampi(i + 1, 0, nx - 1), iw = clampi(i - 1, 0, nx - 1);
    real ht = phb0[I2(j, i)] / G;
    real ht_n = phb0[I2(jn, i)] / G, ht_s = phb0[I2(js, i)] / G;
    real ht_e = phb0[I2(j, ie)] / G, ht_w = phb0[I2(j, iw)] / G;
    real dy_n = ht_n - ht, dy_s = ht - ht_s, dx_e = ht_e - ht, dx_w = ht - ht_w;
    real v_cf_n = (g->cf1 * v[IV(0, j + 1, i)] + g->cf2 * v[IV(1, j + 1, i)]) + g->cf3 * v[IV(2, j + 1, i)];
    real v_cf_s = (g->cf1 * v[IV(0, j, i)] + g->cf2 * v[IV(1, j, i)]) + g->cf3 * v[IV(2, j, i)];
    real u_cf_e = (g->cf1 * u[IU(0, j, i + 1)] + g->cf2 * u[IU(1, j, i + 1)]) + g->cf3 * u[IU(2, j, i + 1)];
    real u_cf_w = (g->cf1 * u[IU(0, j, i)] + g->cf2 * u[IU(1, j, i)]) + g->cf3 * u[IU(2, j, i)];
    return ((msfty * 0.5) * rdy) * (dy_n * v_cf_n + dy_s * v_cf_s) +
           ((msftx * 0.5) * rdx) * (dx_e * u_cf_e + dx_w * u_cf_w);
}

#ifndef ORC_AD
int orc_surface_w(const orc_grid *g, const double *phb0, const double *u, const double *v, double *out) {
    if (!g || !phb0 || !u || !v || !out || g->nz < 3) return -1;
    const int nx = g->nx;
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) out[I2(j, i)] = surface_w(g, phb0, u, v, j, i);
    return 0;
}
#endif

static real *zalloc(size_t n) { return (real *)calloc(n ? n : 1, sizeof(real)); }

/* Extents of the arrays a large step computes on. */
enum { EX_M, EX_W, EX_U, EX_V, EX_C, EX_CU, EX_CV };

static size_t extent_len(int nx, int ny, int nz, int ex) {
    switch (ex) {
        case EX_M: return (size_t)nz * ny * nx;
        case EX_W: return (size_t)(nz + 1) * ny * nx;
        case EX_U: return (size_t)nz * ny * (nx + 1);
        case EX_V: return (size_t)nz * (ny + 1) * nx;
        case EX_C: return (size_t)ny * nx;
        case EX_CU: return (size_t)ny * (nx + 1);
        default: return (size_t)(ny + 1) * nx;
    }
}

/* The state fields, then the workspace arrays, in one fixed order. */
#define N_STATE 11
static const char *const array_names[] = {
    "u", "v", "w", "theta", "p_total", "p_pert", "ph_total", "ph_pert", "mu_total", "mu_pert", "mudf",
    "u0", "v0", "w0", "th0", "ph0", "mu0", "ru", "rv", "rom", "divv", "ut", "vt", "wt", "tht",
    "alt_raw", "al_ls", "php_ls", "mub", "mut", "muts0", "mu_work0", "muu", "muus", "muv", "muvs",
    "u_work", "v_work", "w_work", "th_work", "ph_work", "c2a", "alt", "pb", "phb", "php", "p", "al",
    "pm1", "rw_tend", "ph_tend", "ca", "calpha", "cgamma", "mu", "muts", "muave", "ww", "t2ave",
    "dvdxi", "mudf_new", "ww_new", "wsolve", "rhs", "tmp_ph_total"};
static const int array_extent[] = {
    EX_U, EX_V, EX_W, EX_M, EX_M, EX_M, EX_W, EX_W, EX_C, EX_C, EX_C,
    EX_U, EX_V, EX_W, EX_M, EX_W, EX_C, EX_M, EX_M, EX_W, EX_M, EX_U, EX_V, EX_W, EX_M,
    EX_M, EX_M, EX_M, EX_C, EX_C, EX_C, EX_C, EX_CU, EX_CU, EX_CV, EX_CV,
    EX_U, EX_V, EX_W, EX_M, EX_W, EX_M, EX_M, EX_M, EX_W, EX_M, EX_M, EX_M,
    EX_M, EX_W, EX_W, EX_W, EX_W, EX_W, EX_C, EX_C, EX_C, EX_W, EX_M,
    EX_M, EX_C, EX_W, EX_W, EX_W, EX_W};
#define N_ARRAYS ((int)(sizeof(array_names) / sizeof(array_names[0])))
#define N_WS (N_ARRAYS - N_STATE)

/* Addresses of the state and workspace array pointers, in the order above. */
static void array_slots(rstate *s, ws_t *w, real **slot[N_ARRAYS]) {
    real **all[] = {&s->u, &s->v, &s->w, &s->theta, &s->p_total, &s->p_pert, &s->ph_total,
                    &s->ph_pert, &s->mu_total, &s->mu_pert, &s->mudf,
                    &w->u0, &w->v0, &w->w0, &w->th0, &w->ph0, &w->mu0, &w->ru, &w->rv, &w->rom,
                    &w->divv, &w->ut, &w->vt, &w->wt, &w->tht, &w->alt_raw, &w->al_ls,
                    &w->php_ls, &w->mub, &w->mut, &w->muts0, &w->mu_work0, &w->muu, &w->muus,
                    &w->muv, &w->muvs, &w->u_work, &w->v_work, &w->w_work, &w->th_work,
                    &w->ph_work, &w->c2a, &w->alt, &w->pb, &w->phb, &w->php, &w->p, &w->al,
                    &w->pm1, &w->rw_tend, &w->ph_tend, &w->ca, &w->calpha, &w->cgamma, &w->mu,
                    &w->muts, &w->muave, &w->ww, &w->t2ave, &w->dvdxi, &w->mudf_new,
                    &w->ww_new, &w->wsolve, &w->rhs, &w->tmp_ph_total};
    for (int a = 0; a < N_ARRAYS; ++a) slot[a] = all[a];
}

static int ws_alloc(ws_t *w, int nx, int ny, int nz) {
    rstate unused;
    real **slot[N_ARRAYS];
    memset(w, 0, sizeof(*w));
    array_slots(&unused, w, slot);
    for (int a = N_STATE; a < N_ARRAYS; ++a) {
        *slot[a] = zalloc(extent_len(nx, ny, nz, array_extent[a]));
        if (!*slot[a]) return -1;
    }
    w->mut0 = zalloc((size_t)ny * nx);
    return w->mut0 ? 0 : -1;
}

static void ws_free(ws_t *w) {
    rstate unused;
    real **slot[N_ARRAYS];
    array_slots(&unused, w, slot);
    for (int a = N_STATE; a < N_ARRAYS; ++a) free(*slot[a]);
    free(w->mut0);
}

/* _u_face_average_2d / _v_face_average_2d, core/small_step_prep.py:77-88:
 * edge faces copy the edge cell, interior faces average. */
static void u_face_avg(int nx, int ny, const real *f, real *out) {
    for (int j = 0; j < ny; ++j) {
        out[I2U(j, 0)] = f[I2(j, 0)];
        for (int i = 1; i < nx; ++i) out[I2U(j, i)] = 0.5 * (f[I2(j, i - 1)] + f[I2(j, i)]);
        out[I2U(j, nx)] = f[I2(j, nx - 1)];
    }
}
static void v_face_avg(int nx, int ny, const real *f, real *out) {
    for (int i = 0; i < nx; ++i) {
        out[I2V(0, i)] = f[I2(0, i)];
        for (int j = 1; j < ny; ++j) out[I2V(j, i)] = 0.5 * (f[I2(j - 1, i)] + f[I2(j, i)]);
        out[I2V(ny, i)] = f[I2(ny - 1, i)];
    }
}

/* ------------------------------------------------------------------ */
/* 2-D Smagorinsky horizontal diffusion (diff_opt = 1, km_opt = 4),    */
/* as operational_mode.py:2946-2963 applies it: the eddy viscosity of  */
/* smag2d_horizontal_km (dynamics/explicit_diffusion.py:1116-1217)     */
/* from the deformations of horizontal_deformation_2d (:960-1113),     */
/* called there with no map factors or slopes (so msf = 1, zx = zy =   */
/* 0 and the slope terms vanish), then the coordinate-surface flux     */
/* divergence _hdiff_coord_scalar (:1243-1284) of theta - 300 with     */
/* K / prandtl and of u, v, w with K (:1287-1365).  Periodic in x and  */
/* y over the mass extents; u and v drop their wrap face and take face */
/* 0's tendency there.                                                 */
/* ------------------------------------------------------------------ */
#define SMAG_PRANDTL (1.0 / 3.0)

/* D12 at the south-west corner of mass cell (j, i): du/dy + dv/dx. */
static inline real smag_d12(const orc_grid *g, const real *u, const real *v, int k, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    const real dudy = (u[IU(k, j, i)] - u[IU(k, wrap(j - 1, ny), i)]) / g->dy;
    const real dvdx = (v[IV(k, j, i)] - v[IV(k, j, wrap(i - 1, nx))]) / g->dx;
    return dudy + dvdx;
}

/* xkmh at mass cell (k, j, i): c_s^2 l^2 sqrt(def2), capped at 10 l, with
 * l = sqrt(dx dy) (explicit_diffusion.py:1194-1205). */
static inline real smag_km_at(const orc_grid *g, double c_s, const real *u, const real *v, int k, int j, int i) {
    const int nx = g->nx, ny = g->ny;
    const int ip = wrap(i + 1, nx), jp = wrap(j + 1, ny);
    const real d11 = 2.0 * ((u[IU(k, j, ip)] - u[IU(k, j, i)]) / g->dx);
    const real d22 = 2.0 * ((v[IV(k, jp, i)] - v[IV(k, j, i)]) / g->dy);
    const real tmp = 0.25 * (((smag_d12(g, u, v, k, j, i) + smag_d12(g, u, v, k, jp, i)) +
                              smag_d12(g, u, v, k, j, ip)) + smag_d12(g, u, v, k, jp, ip));
    const real def2 = 0.25 * (d11 - d22) * (d11 - d22) + tmp * tmp;
    const double mlen = sqrt(g->dx * g->dy);
    const real km = c_s * c_s * mlen * mlen * sqrt(def2);
    const double cap = 10.0 * mlen;
    return km > cap ? (real)cap : km;
}

/* rdx (K_e m_e rdx (f_e - f) - K_w m_w rdx (f - f_w)) + the same in y when
 * nrow > 1, at point (j, i) of one level: f, K and mass are given by their
 * values at the centre and the four neighbours (index 0 centre, 1 east,
 * 2 west, 3 north, 4 south). */
static inline real hdiff_point(const real f[5], const real K[5], const real m[5], int nrow, real rdx, real rdy) {
    const real k_e = 0.5 * (K[1] + K[0]), k_w = 0.5 * (K[0] + K[2]);
    const real m_e = 0.5 * (m[1] + m[0]), m_w = 0.5 * (m[0] + m[2]);
    const real p = k_e * m_e * rdx, q = k_w * m_w * rdx;
    real t = rdx * (p * (f[1] - f[0]) - q * (f[0] - f[2]));
    if (nrow > 1) {
        const real k_n = 0.5 * (K[3] + K[0]), k_s = 0.5 * (K[0] + K[4]);
        const real m_n = 0.5 * (m[3] + m[0]), m_s = 0.5 * (m[0] + m[4]);
        const real pn = k_n * m_n * rdy, qs = k_s * m_s * rdy;
        t = t + rdy * (pn * (f[3] - f[0]) - qs * (f[0] - f[4]));
    }
    return t;
}

/* K of the w level f: the mass-level K averaged to the face, the nearest
 * level at faces 0 and nz (_xkmhd_to_w_levels, explicit_diffusion.py:1220-1240). */
static inline real smag_k_w(const real *xkmh, size_t plane, int nz, int f, size_t c) {
    if (f == 0) return xkmh[c];
    if (f == nz) return xkmh[(size_t)(nz - 1) * plane + c];
    return 0.5 * (xkmh[(size_t)f * plane + c] + xkmh[(size_t)(f - 1) * plane + c]);
}

/* The coupled tendencies of the 2-D Smagorinsky diffusion, added to ut,
 * vt, wt and tht. */
static void smagorinsky_hdiff(const orc_grid *g, const orc_namelist *nl, const rstate *S, real *ut, real *vt,
                              real *wt, real *tht) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t plane = (size_t)ny * nx;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    real *K = zalloc((size_t)nz * plane);
    real *muu = zalloc((size_t)ny * (nx + 1)), *muv = zalloc((size_t)(ny + 1) * nx);
    u_face_avg(nx, ny, S->mu_total, muu);
    v_face_avg(nx, ny, S->mu_total, muv);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) K[IM(k, j, i)] = smag_km_at(g, nl->c_s, S->u, S->v, k, j, i);
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int ie = wrap(i + 1, nx), iw = wrap(i - 1, nx), jn = wrap(j + 1, ny), js = wrap(j - 1, ny);
                const int jj[5] = {j, j, j, jn, js}, ii[5] = {i, ie, iw, i, i};
                real f[5], Kp[5], m[5];
                /* w on level k (faces), K averaged to the face, mass c1f mu + c2f */
                for (int n = 0; n < 5; ++n) {
                    const size_t c = (size_t)jj[n] * nx + (size_t)ii[n];
                    f[n] = S->w[(size_t)k * plane + c];
                    Kp[n] = smag_k_w(K, plane, nz, k, c);
                    m[n] = g->c1f[k] * S->mu_total[c] + g->c2f[k];
                }
                wt[IM(k, j, i)] = wt[IM(k, j, i)] + hdiff_point(f, Kp, m, ny, rdx, rdy);
                if (k == nz) continue;
                /* theta - 300 with K / prandtl */
                for (int n = 0; n < 5; ++n) {
                    const size_t c = (size_t)jj[n] * nx + (size_t)ii[n];
                    f[n] = S->theta[(size_t)k * plane + c] - T0;
                    Kp[n] = K[(size_t)k * plane + c] / SMAG_PRANDTL;
                    m[n] = g->c1h[k] * S->mu_total[c] + g->c2h[k];
                }
                tht[IM(k, j, i)] = tht[IM(k, j, i)] + hdiff_point(f, Kp, m, ny, rdx, rdy);
                /* u on face i (faces 0..nx-1, periodic over nx), K of cell i */
                for (int n = 0; n < 5; ++n) {
                    f[n] = S->u[IU(k, jj[n], ii[n])];
                    Kp[n] = K[IM(k, jj[n], ii[n])];
                    m[n] = g->c1h[k] * muu[I2U(jj[n], ii[n])] + g->c2h[k];
                }
                const real du = hdiff_point(f, Kp, m, ny, rdx, rdy);
                ut[IU(k, j, i)] = ut[IU(k, j, i)] + du;
                if (i == 0) ut[IU(k, j, nx)] = ut[IU(k, j, nx)] + du;
                /* v on face j (faces 0..ny-1, periodic over ny), K of cell j */
                for (int n = 0; n < 5; ++n) {
                    f[n] = S->v[IV(k, jj[n], ii[n])];
                    Kp[n] = K[IM(k, jj[n], ii[n])];
                    m[n] = g->c1h[k] * muv[I2V(jj[n], ii[n])] + g->c2h[k];
                }
                const real dv = hdiff_point(f, Kp, m, ny, rdx, rdy);
                vt[IV(k, j, i)] = vt[IV(k, j, i)] + dv;
                if (j == 0) vt[IV(k, ny, i)] = vt[IV(k, ny, i)] + dv;
            }
    free(K);
    free(muu);
    free(muv);
}

/* ------------------------------------------------------------------ */
/* Constant-K diffusion (const_nu_m2_s > 0, operational_mode.py:2878-2929): */
/* conservative_constant_k_diffusion_tendency                           */
/* (dynamics/explicit_diffusion.py:282-331) of u, v, w and theta: the   */
/* divergence of mass K grad f, each face mass the mean of its two      */
/* neighbours, periodic over each array's own extent (u over its nx + 1 */
/* faces, v over its ny + 1), no flux through the first and last level, */
/* and one layer thickness for the whole grid, the mean of              */
/* (ph_total[k+1] - ph_total[k]) / g and at least 1 m                   */
/* (operational_mode.py:2756-2758).                                     */
/* ------------------------------------------------------------------ */
#define CK_PARTS 1024

/* dz in a fixed summation order: each column summed upward, the columns in
 * CK_PARTS interleaved partial sums, the partial sums in order. */
static real constk_dz(const orc_grid *g, const real *ph) {
    const int nz = g->nz;
    const size_t C = (size_t)g->ny * g->nx;
    real *part = zalloc(CK_PARTS);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < CK_PARTS; ++t) {
        real s = 0.0;
        for (size_t c = (size_t)t; c < C; c += CK_PARTS) {
            real col = 0.0;
            for (int k = 0; k < nz; ++k) col = col + (ph[(size_t)(k + 1) * C + c] - ph[(size_t)k * C + c]) / G;
            s = s + col;
        }
        part[t] = s;
    }
    real tot = 0.0;
    for (int t = 0; t < CK_PARTS; ++t) tot = tot + part[t];
    free(part);
    const real mean = tot / (double)((size_t)nz * C);
    return mean < 1.0 ? (real)1.0 : mean; /* NaN stays NaN, as jnp.maximum */
}

/* The diffusion tendency of f (nlev x nrow x ncol) at (k, j, i), with mass
 * c1[k] mf[j, i] + c2[k]. */
static inline real constk_point(const real *f, int nlev, int nrow, int ncol, const double *c1, const double *c2,
                                const real *mf, double K, double dx, double dy, real dz, int k, int j, int i) {
    const size_t P = (size_t)nrow * ncol;
#define CKF(kk, jj, ii) f[(size_t)(kk) * P + (size_t)(jj) * ncol + (size_t)(ii)]
#define CKM(kk, jj, ii) (c1[kk] * mf[(size_t)(jj) * ncol + (size_t)(ii)] + c2[kk])
    const int ie = wrap(i + 1, ncol), iw = wrap(i - 1, ncol);
    const real fe = (0.5 * (CKM(k, j, i) + CKM(k, j, ie))) * K * (CKF(k, j, ie) - CKF(k, j, i)) / dx;
    const real fw = (0.5 * (CKM(k, j, iw) + CKM(k, j, i))) * K * (CKF(k, j, i) - CKF(k, j, iw)) / dx;
    real t = (fe - fw) / dx;
    if (nrow > 1) {
        const int jn = wrap(j + 1, nrow), js = wrap(j - 1, nrow);
        const real fn = (0.5 * (CKM(k, j, i) + CKM(k, jn, i))) * K * (CKF(k, jn, i) - CKF(k, j, i)) / dy;
        const real fs = (0.5 * (CKM(k, js, i) + CKM(k, j, i))) * K * (CKF(k, j, i) - CKF(k, js, i)) / dy;
        t = t + (fn - fs) / dy;
    }
    if (nlev > 2) {
        const real fu = k + 1 < nlev
                            ? (0.5 * (CKM(k, j, i) + CKM(k + 1, j, i))) * K * (CKF(k + 1, j, i) - CKF(k, j, i)) / dz
                            : (real)0.0;
        const real fd = k > 0 ? (0.5 * (CKM(k - 1, j, i) + CKM(k, j, i))) * K * (CKF(k, j, i) - CKF(k - 1, j, i)) / dz
                              : (real)0.0;
        t = t + (fu - fd) / dz;
    }
#undef CKF
#undef CKM
    return t;
}

static void constk_field(const real *f, int nlev, int nrow, int ncol, const double *c1, const double *c2,
                         const real *mf, double K, double dx, double dy, real dz, real *tend) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nlev; ++k)
        for (int j = 0; j < nrow; ++j)
            for (int i = 0; i < ncol; ++i) {
                const size_t q = ((size_t)k * nrow + j) * ncol + i;
                tend[q] = tend[q] + constk_point(f, nlev, nrow, ncol, c1, c2, mf, K, dx, dy, dz, k, j, i);
            }
}

/* The coupled constant-K tendencies added to ut, vt, wt and tht; the layer
 * thickness used in *dz_out when given. */
static void constk_diffusion(const orc_grid *g, double K, const rstate *S, real *ut, real *vt, real *wt, real *tht,
                             real *dz_out) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real dz = constk_dz(g, S->ph_total);
    real *muu = zalloc((size_t)ny * (nx + 1)), *muv = zalloc((size_t)(ny + 1) * nx);
    u_face_avg(nx, ny, S->mu_total, muu);
    v_face_avg(nx, ny, S->mu_total, muv);
    constk_field(S->theta, nz, ny, nx, g->c1h, g->c2h, S->mu_total, K, g->dx, g->dy, dz, tht);
    constk_field(S->u, nz, ny, nx + 1, g->c1h, g->c2h, muu, K, g->dx, g->dy, dz, ut);
    constk_field(S->v, nz, ny + 1, nx, g->c1h, g->c2h, muv, K, g->dx, g->dy, dz, vt);
    constk_field(S->w, nz + 1, ny, nx, g->c1f, g->c2f, S->mu_total, K, g->dx, g->dy, dz, wt);
    if (dz_out) *dz_out = dz;
    free(muu);
    free(muv);
}

#ifndef ORC_AD
int orc_constk_tendency(const orc_grid *g, double nu, const orc_state *s, double *ut, double *vt, double *wt,
                        double *tht, double *dz) {
    if (!g || !s || !ut || !vt || !wt || !tht || g->nx < 1 || g->ny < 1 || g->nz < 1 || !(nu >= 0.0)) return -1;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    memset(ut, 0, (size_t)nz * ny * (nx + 1) * sizeof(double));
    memset(vt, 0, (size_t)nz * (ny + 1) * nx * sizeof(double));
    memset(wt, 0, (size_t)(nz + 1) * ny * nx * sizeof(double));
    memset(tht, 0, (size_t)nz * ny * nx * sizeof(double));
    const rstate S = {s->u, s->v, s->w, s->theta, s->p_total, s->p_pert, s->ph_total, s->ph_pert,
                      s->mu_total, s->mu_pert, s->mudf};
    constk_diffusion(g, nu, &S, ut, vt, wt, tht, dz);
    return 0;
}
#endif

#ifndef ORC_AD
int orc_smag_tendency(const orc_grid *g, double c_s, const orc_state *s, double *ut, double *vt, double *wt,
                      double *tht, double *xkmh) {
    if (!g || !s || !ut || !vt || !wt || !tht || g->nx < 4 || g->ny < 1 || g->nz < 1) return -1;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t M = (size_t)nz * ny * nx;
    memset(ut, 0, (size_t)nz * ny * (nx + 1) * sizeof(double));
    memset(vt, 0, (size_t)nz * (ny + 1) * nx * sizeof(double));
    memset(wt, 0, (size_t)(nz + 1) * ny * nx * sizeof(double));
    memset(tht, 0, M * sizeof(double));
    orc_namelist nl;
    memset(&nl, 0, sizeof nl);
    nl.diff_opt = 1;
    nl.km_opt = 4;
    nl.c_s = c_s;
    const rstate S = {s->u, s->v, s->w, s->theta, s->p_total, s->p_pert, s->ph_total, s->ph_pert,
                      s->mu_total, s->mu_pert, s->mudf};
    smagorinsky_hdiff(g, &nl, &S, ut, vt, wt, tht);
    if (xkmh)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) xkmh[IM(k, j, i)] = smag_km_at(g, c_s, s->u, s->v, k, j, i);
    return 0;
}

/* The moisture coupling factors alone, for tests: qtot, cqu, cqv and the
 * raw cqw of the state's six species. */
int orc_moist_factors(const orc_grid *g, const orc_state *s, double *qtot, double *cqu, double *cqv, double *cqwr) {
    if (!g || !s || !s->qv || !s->qc || !s->qr || !s->qi || !s->qs || !s->qg || g->nx < 1 || g->ny < 1 || g->nz < 2)
        return -1;
    moist_t m;
    memset(&m, 0, sizeof m);
    double *qs[N_MOIST] = {s->qv, s->qc, s->qr, s->qi, s->qs, s->qg};
    for (int sp = 0; sp < N_MOIST; ++sp) m.q[sp] = qs[sp];
    m.qtot = qtot;
    m.cqu = cqu;
    m.cqv = cqv;
    m.cqwr = cqwr;
    moist_factors(g, &m);
    return 0;
}

/* The Coriolis terms alone, for tests: ut, vt get large_step_coriolis and
 * rwt (nz+1 faces) the cosine-Coriolis part of rw_tend, with the face
 * masses of the stage mut = (mu_total - mu_pert) + mu_pert as
 * small_step_prep forms them (0 at faces 0 and nz). */
int orc_coriolis_tendency(const orc_grid *g, const orc_state *s, double *ut, double *vt, double *rwt) {
    if (!g || !s || !ut || !vt || !rwt || g->nx < 4 || g->ny < 1 || g->nz < 2) return -1;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    memset(ut, 0, (size_t)nz * ny * (nx + 1) * sizeof(double));
    memset(vt, 0, (size_t)nz * (ny + 1) * nx * sizeof(double));
    memset(rwt, 0, (size_t)(nz + 1) * ny * nx * sizeof(double));
    const rstate S = {s->u, s->v, s->w, s->theta, s->p_total, s->p_pert, s->ph_total, s->ph_pert,
                      s->mu_total, s->mu_pert, s->mudf};
    large_step_coriolis(g, &S, ut, vt);
    real *mut = zalloc((size_t)ny * nx), *muu = zalloc((size_t)ny * (nx + 1)), *muv = zalloc((size_t)(ny + 1) * nx);
    for (size_t c = 0; c < (size_t)ny * nx; ++c) mut[c] = (s->mu_total[c] - s->mu_pert[c]) + s->mu_pert[c];
    u_face_avg(nx, ny, mut, muu);
    v_face_avg(nx, ny, mut, muv);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real mxy = g->msftx[c] / g->msfty[c];
            real ru_m[2], rv_m[2];
            for (int k = 1; k < nz; ++k) {
                for (int q = 0; q < 2; ++q) {
                    int kk = k - q;
                    real ru_l = ((g->c1h[kk] * muu[I2U(j, i)] + g->c2h[kk]) * s->u[IU(kk, j, i)]) / g->msfuy[I2U(j, i)];
                    real ru_r = ((g->c1h[kk] * muu[I2U(j, i + 1)] + g->c2h[kk]) * s->u[IU(kk, j, i + 1)]) /
                                g->msfuy[I2U(j, i + 1)];
                    ru_m[q] = 0.5 * (ru_l + ru_r);
                    real rv_s = ((g->c1h[kk] * muv[I2V(j, i)] + g->c2h[kk]) * s->v[IV(kk, j, i)]) / g->msfvx[I2V(j, i)];
                    real rv_n = ((g->c1h[kk] * muv[I2V(j + 1, i)] + g->c2h[kk]) * s->v[IV(kk, j + 1, i)]) /
                                g->msfvx[I2V(j + 1, i)];
                    rv_m[q] = 0.5 * (rv_s + rv_n);
                }
                real ru_f = g->fnm[k] * ru_m[0] + g->fnp[k] * ru_m[1];
                real rv_f = g->fnm[k] * rv_m[0] + g->fnp[k] * rv_m[1];
                rwt[IM(k, j, i)] = cor_w(g, c, mxy, ru_f, rv_f);
            }
        }
    free(mut);
    free(muu);
    free(muv);
    return 0;
}
#endif

/* ------------------------------------------------------------------ */
/* Positive-definite (scalar_adv_opt 1) and monotonic (2) limited       */
/* scalar advection, advect_scalar_flux_limited                         */
/* (dynamics/flux_advection.py:614-975), which the reference applies to */
/* theta - 300 on the final RK stage (operational_mode.py:2819-2866):   */
/* donor-cell fluxes from the step-origin field and the antidiffusive  */
/* remainder of the h5/v3 fluxes, scaled per cell so that the explicit */
/* update by dt stays non-negative (1) or within the origin field's    */
/* local extrema (2).  Periodic in x and y, rigid top and bottom.      */
/* ------------------------------------------------------------------ */
#define LIM_EPS 1.0e-20

static inline real lim_pos(real a) { return a < 0.0 ? (real)0.0 : a; }
static inline real lim_neg(real a) { return a > 0.0 ? (real)0.0 : a; }
static inline real lim_min(real a, real b) { return b < a ? b : a; }
static inline real lim_max(real a, real b) { return b > a ? b : a; }

/* Donor-cell flux at the face between q_m (upstream side for cr > 0) and
 * q_p, mass mu_face, spacing dxy (flux_upwind, flux_advection.py:614-640). */
static inline real lim_upwind(real q_m, real q_p, real vel, real mu_face, double dxy, double dt) {
    const real cr = vel * dt / dxy / mu_face;
    const real acr = fabs(cr);
    const real w_m = 0.5 * lim_min(1.0, cr + acr);
    const real w_p = 0.5 * lim_max(-1.0, cr - acr);
    return mu_face * (dxy / dt) * (w_m * q_m + w_p * q_p);
}

/* The scale of a cell's outgoing (PD) or incoming/outgoing (monotonic)
 * antidiffusive flux: 1 unless the flux exceeds the room r. */
static inline real lim_scale(real flux, real room) {
    return flux > room ? lim_max(0.0, room / (flux + LIM_EPS)) : (real)1.0;
}

/* A horizontal face's limited antidiffusive flux a: m the cell on the face's
 * minus side, p the plus side.  PD: the donor cell's scale; monotonic: the
 * smaller of the receiving cell's inflow and the donor's outflow scale. */
static inline real lim_face(int opt, real a, const real *sa, const real *sb, size_t m, size_t p) {
    if (opt == 1) return a > 0.0 ? sa[m] * a : (a < 0.0 ? sa[p] * a : a);
    return a > 0.0 ? lim_min(sa[p], sb[m]) * a : lim_min(sb[p], sa[m]) * a;
}
/* The same at z face f of column (j, i): the mass coordinate decreases with
 * k, so a negative flux drains the cell below (f - 1) and a positive one the
 * cell above (f), each clamped to the column. */
static inline real lim_zface(int opt, int nz, real a, const real *sa, const real *sb, int f, int j, int i, int nx,
                             int ny) {
    const size_t above = IM(f < nz - 1 ? f : nz - 1, j, i), below = IM(f > 0 ? f - 1 : 0, j, i);
    if (opt == 1) return a < 0.0 ? sa[below] * a : (a > 0.0 ? sa[above] * a : a);
    return a < 0.0 ? lim_min(sa[above], sb[below]) * a : lim_min(sb[above], sa[below]) * a;
}

/* tend = the coupled tendency of q (theta - 300 here) with the limiter opt
 * (1 or 2); qold the step-origin field, mut the stage mu_total, mu_old the
 * step-origin mu_total, dt the full step. */
static void advect_scalar_limited(const orc_grid *g, int opt, const real *q, const real *qold, const real *ru,
                                  const real *rv, const real *rom, const real *mut, const real *mu_old, double dt,
                                  real *tend) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t M = (size_t)nz * ny * nx, Wn = (size_t)(nz + 1) * ny * nx;
    const double rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    const double dxl = 1.0 / rdx, dyl = 1.0 / rdy;
    real *fxl = zalloc(M), *fx = zalloc(M), *fyl = zalloc(M), *fy = zalloc(M);
    real *fzl = zalloc(Wn), *fz = zalloc(Wn), *sa = zalloc(M), *sb = zalloc(M);
    /* face fluxes: x face i and y face j of cell (k, j, i) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t c = IM(k, j, i);
                const int im = wrap(i - 1, nx), jm = wrap(j - 1, ny);
                const real muu = 0.5 * (mut[I2(j, i)] + mut[I2(j, im)]);
                const real muv = 0.5 * (mut[I2(j, i)] + mut[I2(jm, i)]);
                const real mu_u = g->c1h[k] * muu + g->c2h[k], mu_v = g->c1h[k] * muv + g->c2h[k];
#define QX(s) q[IM(k, j, wrap(i + (s), nx))]
#define QY(s) q[IM(k, wrap(j + (s), ny), i)]
                fxl[c] = lim_upwind(qold[IM(k, j, im)], qold[c], ru[c], mu_u, dxl, dt);
                fx[c] = ru[c] * flux5(QX(-3), QX(-2), QX(-1), QX(0), QX(1), QX(2), ru[c]) - fxl[c];
                fyl[c] = lim_upwind(qold[IM(k, jm, i)], qold[c], rv[c], mu_v, dyl, dt);
                fy[c] = rv[c] * flux5(QY(-3), QY(-2), QY(-1), QY(0), QY(1), QY(2), rv[c]) - fyl[c];
#undef QX
#undef QY
            }
    /* z faces 1..nz-1 (0 and nz carry no flux) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int f = 1; f < nz; ++f) {
                const size_t c = I2(j, i), o = (size_t)f * ny * nx + c;
                const real mh = g->c1h[f] * mut[c] + g->c2h[f], mh1 = g->c1h[f - 1] * mut[c] + g->c2h[f - 1];
                const real dz = 2.0 / (g->rdnw[f] + g->rdnw[f - 1]);
                const real mu_face = 0.5 * (mh + mh1);
                const real r = rom[o];
                const real cr = r * dt / dz / mu_face;
                const real acr = fabs(cr);
                const real w_m = 0.5 * lim_min(1.0, cr + acr), w_p = 0.5 * lim_max(-1.0, cr - acr);
                fzl[o] = mu_face * (dz / dt) * (w_m * qold[IM(f - 1, j, i)] + w_p * qold[IM(f, j, i)]);
                real hi;
                if (f == 1)
                    hi = r * (g->fnm[1] * q[IM(1, j, i)] + g->fnp[1] * q[IM(0, j, i)]);
                else if (f == nz - 1)
                    hi = r * (g->fnm[nz - 1] * q[IM(nz - 1, j, i)] + g->fnp[nz - 1] * q[IM(nz - 2, j, i)]);
                else
                    hi = r * flux3(q[IM(f - 2, j, i)], q[IM(f - 1, j, i)], q[IM(f, j, i)], q[IM(f + 1, j, i)], -r);
                fz[o] = hi - fzl[o];
            }
    /* per-cell scales: PD sa = scale; monotonic sa = scale_in, sb = scale_out */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t c = IM(k, j, i), cx = IM(k, j, wrap(i + 1, nx)), cy = IM(k, wrap(j + 1, ny), i);
                const size_t zl = c, zu = IM(k + 1, j, i); /* faces k and k+1 */
                const real msf = g->msftx[I2(j, i)];
                const real mass_old = g->c1h[k] * mu_old[I2(j, i)] + g->c2h[k];
                const real div_low = msf * msf * (rdx * (fxl[cx] - fxl[c]) + rdy * (fyl[cy] - fyl[c])) +
                                     msf * g->rdnw[k] * (fzl[zu] - fzl[zl]);
                const real ph_low = mass_old * qold[c] - dt * div_low;
                const real flux_out = dt * (msf * msf * (rdx * (lim_pos(fx[cx]) - lim_neg(fx[c])) +
                                                         rdy * (lim_pos(fy[cy]) - lim_neg(fy[c]))) +
                                            msf * g->rdnw[k] * (lim_neg(fz[zu]) - lim_pos(fz[zl])));
                if (opt == 1) {
                    sa[c] = lim_scale(flux_out, ph_low);
                } else {
                    real qmax = qold[c], qmin = qold[c];
                    const real nb[6] = {qold[IM(k, j, wrap(i - 1, nx))], qold[cx], qold[IM(k, wrap(j - 1, ny), i)],
                                        qold[cy], qold[IM(k > 0 ? k - 1 : 0, j, i)],
                                        qold[IM(k < nz - 1 ? k + 1 : nz - 1, j, i)]};
                    for (int n = 0; n < 6; ++n) {
                        qmax = lim_max(qmax, nb[n]);
                        qmin = lim_min(qmin, nb[n]);
                    }
                    const real flux_in = -dt * (msf * msf * (rdx * (lim_neg(fx[cx]) - lim_pos(fx[c])) +
                                                             rdy * (lim_neg(fy[cy]) - lim_pos(fy[c]))) +
                                                msf * g->rdnw[k] * (lim_pos(fz[zu]) - lim_neg(fz[zl])));
                    const real ph_hi = qmax * mass_old - ph_low, ph_lo = ph_low - qmin * mass_old;
                    sa[c] = lim_scale(flux_in, ph_hi);
                    sb[c] = lim_scale(flux_out, ph_lo);
                }
            }
    /* limited face fluxes plus the donor-cell ones, then the divergence */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int ip = wrap(i + 1, nx), jp = wrap(j + 1, ny);
                const size_t c = IM(k, j, i);
                /* total flux at x face ii of row (k, j), y face jj of column (k, i), z face f of column (j, i) */
#define FXT(ii) (lim_face(opt, fx[IM(k, j, ii)], sa, sb, IM(k, j, wrap((ii) - 1, nx)), IM(k, j, ii)) + fxl[IM(k, j, ii)])
#define FYT(jj) (lim_face(opt, fy[IM(k, jj, i)], sa, sb, IM(k, wrap((jj) - 1, ny), i), IM(k, jj, i)) + fyl[IM(k, jj, i)])
                const real msf = g->msftx[I2(j, i)];
                real t = -msf * rdx * (FXT(ip) - FXT(i));
                t = t - msf * rdy * (FYT(jp) - FYT(j));
                const real fz_lo = lim_zface(opt, nz, fz[c], sa, sb, k, j, i, nx, ny) + fzl[c];
                const real fz_hi = lim_zface(opt, nz, fz[IM(k + 1, j, i)], sa, sb, k + 1, j, i, nx, ny) +
                                   fzl[IM(k + 1, j, i)];
                tend[c] = t - g->rdnw[k] * (fz_hi - fz_lo);
#undef FXT
#undef FYT
            }
    free(fxl); free(fx); free(fyl); free(fy); free(fzl); free(fz); free(sa); free(sb);
}

#ifndef ORC_AD
int orc_advect_limited(const orc_grid *g, int opt, const double *u, const double *v, const double *mu_total,
                       const double *q, const double *qold, const double *mu_old, double dt, double *tend) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (opt != 1 && opt != 2) return -1;
    size_t m = (size_t)nz * ny * nx, w = (size_t)(nz + 1) * ny * nx;
    real *ru = zalloc(m), *rv = zalloc(m), *rom = zalloc(w), *dv = zalloc(m);
    couple_velocities(g, u, v, mu_total, ru, rv, rom, dv);
    advect_scalar_limited(g, opt, q, qold, ru, rv, rom, mu_total, mu_old, dt, tend);
    free(ru); free(rv); free(rom); free(dv);
    return 0;
}
#endif

/* ------------------------------------------------------------------ */
/* One RK stage: operational_mode.py:3331-3515 advance_stage, as its    */
/* pieces in order: stage_setup, then nsub times (substep_uv,           */
/* substep_mu_theta, substep_w, substep_p), then stage_finish.  Each    */
/* piece reads and writes only the state S and the workspace W.         */
/* ------------------------------------------------------------------ */
static void stage_setup(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W, int rk_step,
                        double dts) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    const size_t Mn = (size_t)nz * ny * nx;

    /* --- transport velocities (operational_mode.py:2429-2475) --- */
    couple_velocities(g, S->u, S->v, S->mu_total, W->ru, W->rv, W->rom, W->divv);

    /* --- large-step tendencies (operational_mode.py:2722-3093) --- */
    advect_u_flux(g, S->u, W->ru, W->rv, W->rom, W->ut);
    advect_v_flux(g, S->v, W->ru, W->rv, W->rom, W->vt);
    advect_w_flux(g, S->w, W->ru, W->rv, W->rom, W->wt);
    /* theta - 300 advected scalar */
    {
        real *thp = W->th_work; /* temporary use before prep overwrites it */
        for (size_t q = 0; q < Mn; ++q) thp[q] = S->theta[q] - T0;
        if (nl->scalar_adv_opt != 0 && rk_step == 3) {
            /* the limiter on the final stage, from the step origin
             * (operational_mode.py:2819-2851) */
            real *tho = zalloc(Mn);
            for (size_t q = 0; q < Mn; ++q) tho[q] = W->th0[q] - T0;
            advect_scalar_limited(g, nl->scalar_adv_opt, thp, tho, W->ru, W->rv, W->rom, S->mu_total, W->mut0,
                                  nl->dt, W->tht);
            free(tho);
        } else {
            advect_scalar_flux(g, thp, W->ru, W->rv, W->rom, W->tht);
        }
    }
    /* the moisture's coupling factors and, with moist_adv_opt, its coupled
     * tendencies with the same transport velocities: the plain h5/v3 flux
     * advection, the limiter on the final stage from the step origin
     * (operational_mode.py:3162-3230, flux_advection.py:1002-1099) */
    if (W->moist) {
        moist_t *m = W->moist;
        moist_factors(g, m);
        if (nl->moist_adv_opt != 0)
            for (int sp = 0; sp < N_MOIST; ++sp) {
                if (rk_step == 3)
                    advect_scalar_limited(g, nl->moist_adv_opt, m->q[sp], m->q0[sp], W->ru, W->rv, W->rom,
                                          S->mu_total, W->mut0, nl->dt, m->qt[sp]);
                else
                    advect_scalar_flux(g, m->q[sp], W->ru, W->rv, W->rom, m->qt[sp]);
            }
    }
    /* 6th-order filter on the coupled u, v, w, theta tendencies, with the
     * face masses of the stage mu_total (operational_mode.py:2768-2774,
     * 2870-2876), before the large-step PGF is added. */
    if (nl->diff_6th_opt != 0) {
        const double coef = nl->diff_6th_factor * 0.015625 / (2.0 * nl->dt);
        real *muu6 = zalloc((size_t)ny * (nx + 1)), *muv6 = zalloc((size_t)(ny + 1) * nx);
        u_face_avg(nx, ny, S->mu_total, muu6);
        v_face_avg(nx, ny, S->mu_total, muv6);
        sixth_filter(S->u, nz, ny, nx + 1, coef, g->c1h, g->c2h, muu6, W->ut);
        sixth_filter(S->v, nz, ny + 1, nx, coef, g->c1h, g->c2h, muv6, W->vt);
        sixth_filter(S->w, nz + 1, ny, nx, coef, g->c1f, g->c2f, S->mu_total, W->wt);
        sixth_filter(S->theta, nz, ny, nx, coef, g->c1h, g->c2h, S->mu_total, W->tht);
        free(muu6);
        free(muv6);
    }
    /* 2-D Smagorinsky horizontal diffusion (operational_mode.py:2946-2963), added
     * after the filter and before the PGF. */
    /* constant-K diffusion (operational_mode.py:2878-2929), after the filter */
    if (nl->const_nu_m2_s > 0.0) constk_diffusion(g, nl->const_nu_m2_s, S, W->ut, W->vt, W->wt, W->tht, NULL);
    if (nl->diff_opt == 1 && nl->km_opt == 4) smagorinsky_hdiff(g, nl, S, W->ut, W->vt, W->wt, W->tht);
    /* EOS alt of the stage-entry state, shared by _absolute_diagnostics
     * (rk_addtend_dry.py:184-186) and diagnose_pressure_al_alt's base_state=None
     * branch used by small_step_prep (acoustic_wrf.py:354-358). */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < Mn; ++q) W->alt_raw[q] = alt_from_theta_p(S->theta[q], S->p_total[q]);
    large_step_pgf(g, nl->hypsometric_opt, S, W->alt_raw, W->al_ls, W->php_ls, W->ut, W->vt, NULL,
                   W->moist ? W->moist->cqu : NULL, W->moist ? W->moist->cqv : NULL);
    large_step_coriolis(g, S, W->ut, W->vt);

    /* --- small_step_prep_wrf, core/small_step_prep.py:216-394 --- */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real mub = S->mu_total[c] - S->mu_pert[c];
            real mu_cur = S->mu_pert[c];
            real mu_ref = W->mu0[c];
            real mu_work = rk_step == 1 ? 0.0 : mu_ref - mu_cur;
            real mut = mub + mu_cur;
            W->mub[c] = mub;
            W->mut[c] = mut;
            W->mu_work0[c] = mu_work;
            W->muts0[c] = mut + mu_work;
            W->mu[c] = mu_cur + mu_work; /* acoustic mu = mu_save + mu_work */
        }
    u_face_avg(nx, ny, W->mut, W->muu);
    v_face_avg(nx, ny, W->mut, W->muv);
    u_face_avg(nx, ny, W->muts0, W->muus);
    v_face_avg(nx, ny, W->muts0, W->muvs);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i <= nx; ++i) {
                real mref = g->c1h[k] * W->muus[I2U(j, i)] + g->c2h[k];
                real mcur = g->c1h[k] * W->muu[I2U(j, i)] + g->c2h[k];
                W->u_work[IU(k, j, i)] =
                    (mref * W->u0[IU(k, j, i)] - mcur * S->u[IU(k, j, i)]) / g->msfuy[I2U(j, i)];
            }
        }
        for (int j = 0; j <= ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                real mref = g->c1h[k] * W->muvs[I2V(j, i)] + g->c2h[k];
                real mcur = g->c1h[k] * W->muv[I2V(j, i)] + g->c2h[k];
                W->v_work[IV(k, j, i)] =
                    (mref * W->v0[IV(k, j, i)] - mcur * S->v[IV(k, j, i)]) / g->msfvx[I2V(j, i)];
            }
        }
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real mref = g->c1h[k] * W->muts0[c] + g->c2h[k];
                real mcur = g->c1h[k] * W->mut[c] + g->c2h[k];
                real th_ref = W->th0[q] - T0, th_cur = S->theta[q] - T0;
                W->th_work[q] = mref * th_ref - mcur * th_cur;
                real pb = S->p_total[q] - S->p_pert[q];
                W->pb[q] = pb;
                real alt_ac = W->alt_raw[q] > MIN_ACOUSTIC_ALT ? W->alt_raw[q] : MIN_ACOUSTIC_ALT;
                W->alt[q] = alt_ac;
                W->c2a[q] = (CPOVCV * (pb + S->p_pert[q])) / alt_ac;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real mref = g->c1f[k] * W->muts0[c] + g->c2f[k];
                real mcur = g->c1f[k] * W->mut[c] + g->c2f[k];
                W->w_work[q] = (mref * W->w0[q] - mcur * S->w[q]) / g->msfty[c];
                W->ph_work[q] = W->ph0[q] - S->ph_pert[q];
                W->phb[q] = S->ph_total[q] - S->ph_pert[q];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t q = IM(k, j, i), q1 = IM(k + 1, j, i);
                W->php[q] = 0.5 * (((W->phb[q] + W->phb[q1]) + S->ph_pert[q]) + S->ph_pert[q1]);
            }

    /* --- calc_p_rho_wrf step 0, core/calc_p_rho.py:48-139 --- */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real mass_h = g->c1h[k] * W->muts0[c] + g->c2h[k];
                real safe_mass = safe_pos(mass_h, 1.0e-12);
                real mu_term = g->c1h[k] * W->mu_work0[c];
                real al = -(W->alt[q] * mu_term + g->rdnw[k] * (W->ph_work[IM(k + 1, j, i)] - W->ph_work[q])) /
                            safe_mass;
                real theta_1 = W->th0[q] - T0;
                real ttr = T0 + theta_1;
                real safe_t = safe_pos(ttr, 1.0e-6);
                real p = W->c2a[q] * (((W->alt[q] * (W->th_work[q] - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
                W->al[q] = al;
                W->p[q] = p;
                W->pm1[q] = p;
            }

    /* --- _acoustic_core_state_from_prep, operational_mode.py:1682-2166 --- */
    /* rw_tend: pg_buoy_w_moist with qtot=0 (core/advance_w.py:209-290) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real msft_inv = 1.0 / g->msfty[c];
            real mu_p = W->mut[c] - W->mub[c];
            real mub = W->mub[c];
            W->rw_tend[IM(0, j, i)] = 0.0;
            for (int k = 1; k < nz; ++k) {
                real raw = W->moist ? W->moist->cqwr[IM(k, j, i)] : (real)0.0;
                real cq1 = 1.0 / (1.0 + raw), cq2 = raw * cq1;
                real rw = (msft_inv * G) *
                            (((cq1 * g->rdn[k]) * (S->p_pert[IM(k, j, i)] - S->p_pert[IM(k - 1, j, i)]) -
                              g->c1f[k] * mu_p) -
                             cq2 * (g->c1f[k] * mub + g->c2f[k]));
                W->rw_tend[IM(k, j, i)] = rw;
            }
            {
                /* the top face takes the raw cqw of face nz - 1 */
                real raw = W->moist ? W->moist->cqwr[IM(nz - 1, j, i)] : (real)0.0;
                real cq1 = 1.0 / (1.0 + raw), cq2 = raw * cq1;
                real rw = (msft_inv * G) *
                            (((((cq1 * 2.0) * g->rdnw[nz - 1]) * (-S->p_pert[IM(nz - 1, j, i)])) -
                              g->c1f[nz] * mu_p) -
                             cq2 * (g->c1f[nz] * mub + g->c2f[nz]));
                W->rw_tend[IM(nz, j, i)] = rw;
            }
        }
    /* + tendencies.w (operational_mode.py:1782) and curvature (:1793-1822) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            for (int k = 0; k <= nz; ++k) W->rw_tend[IM(k, j, i)] = W->rw_tend[IM(k, j, i)] + W->wt[IM(k, j, i)];
            real mxy = g->msftx[c] / g->msfty[c];
            real ru_m[2], u_m[2], rv_m[2], v_m[2];
            for (int k = 1; k < nz; ++k) {
                for (int s = 0; s < 2; ++s) {
                    int kk = k - s; /* s=0: level k, s=1: level k-1 */
                    real mu_l = g->c1h[kk] * W->muu[I2U(j, i)] + g->c2h[kk];
                    real mu_r = g->c1h[kk] * W->muu[I2U(j, i + 1)] + g->c2h[kk];
                    real ru_l = (mu_l * S->u[IU(kk, j, i)]) / g->msfuy[I2U(j, i)];
                    real ru_r = (mu_r * S->u[IU(kk, j, i + 1)]) / g->msfuy[I2U(j, i + 1)];
                    ru_m[s] = 0.5 * (ru_l + ru_r);
                    u_m[s] = 0.5 * (S->u[IU(kk, j, i)] + S->u[IU(kk, j, i + 1)]);
                    real mv_s = g->c1h[kk] * W->muv[I2V(j, i)] + g->c2h[kk];
                    real mv_n = g->c1h[kk] * W->muv[I2V(j + 1, i)] + g->c2h[kk];
                    real rv_s = (mv_s * S->v[IV(kk, j, i)]) / g->msfvx[I2V(j, i)];
                    real rv_n = (mv_n * S->v[IV(kk, j + 1, i)]) / g->msfvx[I2V(j + 1, i)];
                    rv_m[s] = 0.5 * (rv_s + rv_n);
                    v_m[s] = 0.5 * (S->v[IV(kk, j, i)] + S->v[IV(kk, j + 1, i)]);
                }
                real ru_f = g->fnm[k] * ru_m[0] + g->fnp[k] * ru_m[1];
                real u_f = g->fnm[k] * u_m[0] + g->fnp[k] * u_m[1];
                real rv_f = g->fnm[k] * rv_m[0] + g->fnp[k] * rv_m[1];
                real v_f = g->fnm[k] * v_m[0] + g->fnp[k] * v_m[1];
                real curv = RERADIUS * (ru_f * u_f + (mxy * rv_f) * v_f);
                W->rw_tend[IM(k, j, i)] = W->rw_tend[IM(k, j, i)] + (cor_w(g, c, mxy, ru_f, rv_f) + curv);
            }
        }
    /* rhs_ph_wrf order-2 periodic (core/rhs_ph.py:175-398), top_lid */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real msft_inv = 1.0 / g->msfty[c];
            int jp = wrap(j + 1, ny), jm = wrap(j - 1, ny), ip = wrap(i + 1, nx), im = wrap(i - 1, nx);
#define PHT(k, jj, ii) (S->ph_pert[IM((k), (jj), (ii))] + W->phb[IM((k), (jj), (ii))])
            W->ph_tend[IM(0, j, i)] = 0.0;
            for (int f = 1; f < nz; ++f) {
                real wd_f = ((0.5 * (W->rom[IM(f + 1, j, i)] + W->rom[IM(f, j, i)])) * g->rdnw[f]) *
                              (PHT(f + 1, j, i) - PHT(f, j, i));
                real wd_fm = ((0.5 * (W->rom[IM(f, j, i)] + W->rom[IM(f - 1, j, i)])) * g->rdnw[f - 1]) *
                               (PHT(f, j, i) - PHT(f - 1, j, i));
                real term3 = g->fnm[f] * wd_f + g->fnp[f] * wd_fm;
                real t = -term3;
                real mass_f = g->c1f[f] * W->mut[c] + g->c2f[f];
                real gw = ((mass_f * G) * S->w[IM(f, j, i)]) * msft_inv;
                t = t + gw;
                /* y advection */
                real pti = PHT(f, j, i);
                real dn_ = PHT(f, jp, i) - pti, ds_ = pti - PHT(f, jm, i);
                real vn = S->v[IV(f, j + 1, i)] + S->v[IV(f - 1, j + 1, i)];
                real vs = S->v[IV(f, j, i)] + S->v[IV(f - 1, j, i)];
                real mvn = g->c1f[f] * W->muv[I2V(j + 1, i)] + g->c2f[f];
                real mvs = g->c1f[f] * W->muv[I2V(j, i)] + g->c2f[f];
                real fy = (mvn * vn) * dn_ + (mvs * vs) * ds_;
                t = t + (-((0.25 * rdy) * msft_inv)) * fy;
                real de_ = PHT(f, j, ip) - pti, dw_ = pti - PHT(f, j, im);
                real ue = S->u[IU(f, j, i + 1)] + S->u[IU(f - 1, j, i + 1)];
                real uw = S->u[IU(f, j, i)] + S->u[IU(f - 1, j, i)];
                real mue = g->c1f[f] * W->muu[I2U(j, i + 1)] + g->c2f[f];
                real muw = g->c1f[f] * W->muu[I2U(j, i)] + g->c2f[f];
                real fx = (mue * ue) * de_ + (muw * uw) * dw_;
                t = t + (-((0.25 * rdx) * msft_inv)) * fx;
                W->ph_tend[IM(f, j, i)] = t;
            }
            W->ph_tend[IM(nz, j, i)] = 0.0;
#undef PHT
        }
    /* calc_coef_w with the stage mut and c2a (operational_mode.py:2346-2354) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            calc_coef_w_column(g, W->mut[I2(j, i)], W->c2a + I2(j, i), (size_t)ny * nx, dts, nl->epssm,
                               nl->top_lid, W->ca + I2(j, i), W->calpha + I2(j, i), W->cgamma + I2(j, i),
                               (size_t)ny * nx, W->moist ? W->moist->cqwr + I2(j, i) : NULL, (size_t)ny * nx);

    /* acoustic carried state initialisation */
    memcpy(W->muts, W->muts0, sizeof(real) * (size_t)ny * nx);
    memset(W->muave, 0, sizeof(real) * (size_t)ny * nx);
    memcpy(W->ww, W->rom, sizeof(real) * (size_t)(nz + 1) * ny * nx);
    memset(W->t2ave, 0, sizeof(real) * Mn);
}

static void substep_uv(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W, double dts) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    const real *mudf = S->mudf;
    (void)nl;

    /* ---- advance_uv_wrf, core/acoustic.py:626-786 (emdiv=0.01) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        real *dpn = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        for (int i = 0; i <= nx; ++i) {
            int il = clampi(i - 1, 0, nx - 1), ir = clampi(i, 0, nx - 1);
            real msf = g->msfux[I2U(j, i)] / g->msfuy[I2U(j, i)];
            real muw_l = W->muts[I2(j, il)] - W->mut[I2(j, il)];
            real muw_r = W->muts[I2(j, ir)] - W->mut[I2(j, ir)];
            {
                real ps0 = W->p[IM(0, j, il)] + W->p[IM(0, j, ir)];
                real ps1 = W->p[IM(1, j, il)] + W->p[IM(1, j, ir)];
                real ps2 = W->p[IM(2, j, il)] + W->p[IM(2, j, ir)];
                dpn[0] = 0.5 * ((g->cf1 * ps0 + g->cf2 * ps1) + g->cf3 * ps2);
                for (int k = 1; k < nz; ++k) {
                    real psk = W->p[IM(k, j, il)] + W->p[IM(k, j, ir)];
                    real pskm = W->p[IM(k - 1, j, il)] + W->p[IM(k - 1, j, ir)];
                    dpn[k] = 0.5 * (g->fnm[k] * psk + g->fnp[k] * pskm);
                }
                real pt = W->p[IM(nz - 1, j, il)] + W->p[IM(nz - 1, j, ir)];
                real pt2 = W->p[IM(nz - 2, j, il)] + W->p[IM(nz - 2, j, ir)];
                real pt3 = W->p[IM(nz - 3, j, il)] + W->p[IM(nz - 3, j, ir)];
                dpn[nz] = 0.5 * ((g->cf1 * pt + g->cf2 * pt2) + g->cf3 * pt3);
            }
            real mudf_xy = ((-EMDIV * g->dx) * (mudf[I2(j, ir)] - mudf[I2(j, il)])) / g->msfuy[I2U(j, i)];
            for (int k = 0; k < nz; ++k) {
                size_t q = IU(k, j, i);
                real u = W->u_work[q] + dts * W->ut[q];
                real mass_x = g->c1h[k] * W->muu[I2U(j, i)] + g->c2h[k];
                real ph_term = (W->ph_work[IM(k + 1, j, ir)] - W->ph_work[IM(k + 1, j, il)]) +
                                 (W->ph_work[IM(k, j, ir)] - W->ph_work[IM(k, j, il)]);
                real p_term = (W->alt[IM(k, j, il)] + W->alt[IM(k, j, ir)]) *
                                (W->p[IM(k, j, ir)] - W->p[IM(k, j, il)]);
                real pb_term = (W->al[IM(k, j, il)] + W->al[IM(k, j, ir)]) *
                                 (W->pb[IM(k, j, ir)] - W->pb[IM(k, j, il)]);
                real dpx = (((msf * 0.5) * rdx) * mass_x) * ((ph_term + p_term) + pb_term);
                real bracket = g->rdnw[k] * (dpn[k + 1] - dpn[k]) - 0.5 * (g->c1h[k] * (muw_l + muw_r));
                dpx = dpx + ((msf * rdx) * (W->php[IM(k, j, ir)] - W->php[IM(k, j, il)])) * bracket;
                u = u - (dts * (W->moist ? W->moist->cqu[q] : (real)1.0)) * dpx;
                u = u + g->c1h[k] * mudf_xy;
                W->u_work[q] = u;
            }
        }
        free(dpn);
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny; ++j) {
        real *dpn = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        int js = clampi(j - 1, 0, ny - 1), jn = clampi(j, 0, ny - 1);
        for (int i = 0; i < nx; ++i) {
            real msf = g->msfvy[I2V(j, i)] / g->msfvx[I2V(j, i)];
            real muw_s = W->muts[I2(js, i)] - W->mut[I2(js, i)];
            real muw_n = W->muts[I2(jn, i)] - W->mut[I2(jn, i)];
            {
                real ps0 = W->p[IM(0, js, i)] + W->p[IM(0, jn, i)];
                real ps1 = W->p[IM(1, js, i)] + W->p[IM(1, jn, i)];
                real ps2 = W->p[IM(2, js, i)] + W->p[IM(2, jn, i)];
                dpn[0] = 0.5 * ((g->cf1 * ps0 + g->cf2 * ps1) + g->cf3 * ps2);
                for (int k = 1; k < nz; ++k) {
                    real psk = W->p[IM(k, js, i)] + W->p[IM(k, jn, i)];
                    real pskm = W->p[IM(k - 1, js, i)] + W->p[IM(k - 1, jn, i)];
                    dpn[k] = 0.5 * (g->fnm[k] * psk + g->fnp[k] * pskm);
                }
                real pt = W->p[IM(nz - 1, js, i)] + W->p[IM(nz - 1, jn, i)];
                real pt2 = W->p[IM(nz - 2, js, i)] + W->p[IM(nz - 2, jn, i)];
                real pt3 = W->p[IM(nz - 3, js, i)] + W->p[IM(nz - 3, jn, i)];
                dpn[nz] = 0.5 * ((g->cf1 * pt + g->cf2 * pt2) + g->cf3 * pt3);
            }
            real msfvx_inv = 1.0 / g->msfvx[I2V(j, i)];
            real mudf_xy = ((-EMDIV * g->dy) * (mudf[I2(jn, i)] - mudf[I2(js, i)])) * msfvx_inv;
            for (int k = 0; k < nz; ++k) {
                size_t q = IV(k, j, i);
                real v = W->v_work[q] + dts * W->vt[q];
                real mass_y = g->c1h[k] * W->muv[I2V(j, i)] + g->c2h[k];
                real ph_term = (W->ph_work[IM(k + 1, jn, i)] - W->ph_work[IM(k + 1, js, i)]) +
                                 (W->ph_work[IM(k, jn, i)] - W->ph_work[IM(k, js, i)]);
                real p_term = (W->alt[IM(k, js, i)] + W->alt[IM(k, jn, i)]) *
                                (W->p[IM(k, jn, i)] - W->p[IM(k, js, i)]);
                real pb_term = (W->al[IM(k, js, i)] + W->al[IM(k, jn, i)]) *
                                 (W->pb[IM(k, jn, i)] - W->pb[IM(k, js, i)]);
                real dpy = (((msf * 0.5) * rdy) * mass_y) * ((ph_term + p_term) + pb_term);
                real bracket = g->rdnw[k] * (dpn[k + 1] - dpn[k]) - 0.5 * (g->c1h[k] * (muw_s + muw_n));
                dpy = dpy + ((msf * rdy) * (W->php[IM(k, jn, i)] - W->php[IM(k, js, i)])) * bracket;
                v = v - (dts * (W->moist ? W->moist->cqv[q] : (real)1.0)) * dpy;
                v = v + g->c1h[k] * mudf_xy;
                W->v_work[q] = v;
            }
        }
        free(dpn);
    }
}

static void substep_mu_theta(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W, double dts) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real rdx = 1.0 / g->dx, rdy = 1.0 / g->dy;
    const real eps_p = 1.0 + nl->epssm, eps_m = 1.0 - nl->epssm;
    real *mudf = S->mudf;
    (void)nl;

    /* ---- advance_mu_t_wrf periodic, dynamics/mu_t_advance.py:139-292 ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real msftx = g->msftx[c], msfty = g->msfty[c];
            real dmdt = 0.0;
            for (int k = 0; k < nz; ++k) {
                real c1 = g->c1h[k], c2 = g->c2h[k];
                real v_n = W->v_work[IV(k, j + 1, i)] +
                             ((c1 * W->muv[I2V(j + 1, i)] + c2) * S->v[IV(k, j + 1, i)]) *
                                 (1.0 / g->msfvx[I2V(j + 1, i)]);
                real v_s = W->v_work[IV(k, j, i)] +
                             ((c1 * W->muv[I2V(j, i)] + c2) * S->v[IV(k, j, i)]) * (1.0 / g->msfvx[I2V(j, i)]);
                real u_e = W->u_work[IU(k, j, i + 1)] +
                             ((c1 * W->muu[I2U(j, i + 1)] + c2) * S->u[IU(k, j, i + 1)]) / g->msfuy[I2U(j, i + 1)];
                real u_w = W->u_work[IU(k, j, i)] +
                             ((c1 * W->muu[I2U(j, i)] + c2) * S->u[IU(k, j, i)]) / g->msfuy[I2U(j, i)];
                real d = (msftx * msfty) * (rdy * (v_n - v_s) + rdx * (u_e - u_w));
                W->dvdxi[IM(k, j, i)] = d;
                dmdt = (k == 0) ? g->dnw[0] * d : dmdt + g->dnw[k] * d;
            }
            real mu_work_old = W->muts[c] - W->mut[c];
            real mu_save = W->mu[c] - mu_work_old;
            real mu_tendency = dmdt + 0.0;
            real mu_work_new = mu_work_old + dts * mu_tendency;
            W->mu[c] = mu_save + mu_work_new;
            W->mudf_new[c] = mu_tendency;
            W->muts[c] = W->mut[c] + mu_work_new;
            W->muave[c] = 0.5 * (eps_p * mu_work_new + eps_m * mu_work_old);
            /* ww recurrence */
            real wprev = W->ww[IM(0, j, i)];
            W->ww_new[IM(0, j, i)] = wprev - W->rom[IM(0, j, i)];
            for (int kk = 1; kk < nz; ++kk) {
                int k = kk - 1;
                real incr = (g->dnw[kk - 1] *
                               ((g->c1h[k] * dmdt + W->dvdxi[IM(kk - 1, j, i)]) + g->c1h[k] * 0.0)) /
                              msfty;
                real wk = wprev - incr;
                W->ww_new[IM(kk, j, i)] = wk - W->rom[IM(kk, j, i)];
                wprev = wk;
            }
            W->ww_new[IM(nz, j, i)] = W->ww[IM(nz, j, i)];
        }
    /* theta update (uses theta_1 = t_save neighbours, ww_updated) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            int ip = wrap(i + 1, nx), im = wrap(i - 1, nx), jp = wrap(j + 1, ny), jm = wrap(j - 1, ny);
            real msftx = g->msftx[c], msfty = g->msfty[c];
#define T1(k, jj, ii) (S->theta[IM((k), (jj), (ii))] - T0)
            real wd_lo = 0.0;
            for (int k = 0; k < nz; ++k) {
                real wd_hi;
                if (k + 1 < nz) {
                    int f = k + 1;
                    real face = g->fnm[f] * T1(f, j, i) + g->fnp[f] * T1(f - 1, j, i);
                    wd_hi = W->ww_new[IM(f, j, i)] * face;
                } else {
                    wd_hi = 0.0;
                }
                real th = T1(k, j, i);
                real v_flux = W->v_work[IV(k, j + 1, i)] * (T1(k, jp, i) + th) -
                                W->v_work[IV(k, j, i)] * (th + T1(k, jm, i));
                real u_flux = W->u_work[IU(k, j, i + 1)] * (T1(k, j, ip) + th) -
                                W->u_work[IU(k, j, i)] * (th + T1(k, j, im));
                real tendency = msftx * ((0.5 * rdy) * v_flux + (0.5 * rdx) * u_flux) +
                                  g->rdnw[k] * (wd_hi - wd_lo);
                size_t q = IM(k, j, i);
                real theta_i = W->th_work[q] + ((msfty * dts) * W->tht[q]);
                W->th_work[q] = theta_i - ((dts * msfty) * tendency);
                wd_lo = wd_hi;
            }
#undef T1
        }
    memcpy(W->ww, W->ww_new, sizeof(real) * (size_t)(nz + 1) * ny * nx);
    memcpy(mudf, W->mudf_new, sizeof(real) * (size_t)ny * nx);
}

static void substep_w(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W, double dts) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const real eps_p = 1.0 + nl->epssm, eps_m = 1.0 - nl->epssm;

    /* ---- advance_w_wrf, core/advance_w.py:293-627 ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        real *rhs = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        real *wn = (real *)malloc((size_t)(nz + 1) * sizeof(real));
        real *t2n = (real *)malloc((size_t)nz * sizeof(real));
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real msfty = g->msfty[c];
            real msft_inv = 1.0 / msfty;
            real mut = W->mut[c], muts = W->muts[c], muave = W->muave[c];
#define WQ(k) W->w_work[IM((k), j, i)]
            /* w_damp_vertical_cfl on rw_tend (interior faces), into a local copy */
            /* t_2ave */
            for (int k = 0; k < nz; ++k) {
                size_t q = IM(k, j, i);
                real mass_h = g->c1h[k] * muts + g->c2h[k];
                real half = 0.5 * (eps_p * W->th_work[q] + eps_m * W->t2ave[q]);
                real ttr = T0 + (S->theta[q] - T0);
                t2n[k] = (half + (g->c1h[k] * muave) * T0) / (mass_h * ttr);
            }
            /* rhs of the phi equation */
            rhs[0] = 0.0;
            const real c05geps = (0.5 * G) * eps_m;
            for (int f = 1; f <= nz; ++f) rhs[f] = dts * (W->ph_tend[IM(f, j, i)] + c05geps * WQ(f));
            for (int f = 1; f < nz; ++f) {
#define PH1T(k) (W->ph0[IM((k), j, i)] + W->phb[IM((k), j, i)])
                real wd_f1 = ((0.5 * (W->ww[IM(f + 1, j, i)] + W->ww[IM(f, j, i)])) * g->rdnw[f]) *
                               (PH1T(f + 1) - PH1T(f));
                real wd_f = ((0.5 * (W->ww[IM(f, j, i)] + W->ww[IM(f - 1, j, i)])) * g->rdnw[f - 1]) *
                              (PH1T(f) - PH1T(f - 1));
                real adv = dts * (g->fnm[f] * wd_f1 + g->fnp[f] * wd_f);
                rhs[f] = rhs[f] + (-adv);
            }
            for (int f = 1; f <= nz; ++f) {
                real mass_f = g->c1f[f] * mut + g->c2f[f];
                rhs[f] = W->ph_work[IM(f, j, i)] + (msfty * rhs[f]) / mass_f;
            }
            if (nl->top_lid) rhs[nz] = 0.0;
            /* surface w from terrain */
            real w_surface = surface_w(g, W->phb, S->u, S->v, j, i);
            /* explicit w */
            for (int f = 0; f <= nz; ++f) {
                real rw = W->rw_tend[IM(f, j, i)];
                if (nl->w_damping == 1 && f >= 1 && f < nz) {
                    real mass_f = g->c1f[f] * mut + g->c2f[f];
                    real safe_mf = safe_pos(mass_f, 1.0e-12);
                    real vert_cfl = fabs(((W->ww[IM(f, j, i)] / safe_mf) * g->rdnw[f]) * dts);
                    real damp = (((-sgn(WQ(f))) * W_ALPHA) * (vert_cfl - W_BETA)) * mass_f;
                    rw = rw + (vert_cfl > W_BETA ? damp : 0.0);
                }
                wn[f] = WQ(f) + dts * rw;
            }
            const real c05dtg = (0.5 * dts) * G;
            for (int f = 1; f < nz; ++f) {
                real mh_u = safe_pos(g->c1h[f] * mut + g->c2h[f], 1.0e-12);
                real mh_l = safe_pos(g->c1h[f - 1] * mut + g->c2h[f - 1], 1.0e-12);
                real coef_u = (W->c2a[IM(f, j, i)] * g->rdnw[f]) / mh_u;
                real coef_l = (W->c2a[IM(f - 1, j, i)] * g->rdnw[f - 1]) / mh_l;
                real phk1 = W->ph_work[IM(f + 1, j, i)], phk = W->ph_work[IM(f, j, i)],
                       phkm = W->ph_work[IM(f - 1, j, i)];
                real tA_u = coef_u * (eps_p * (rhs[f + 1] - rhs[f]) + eps_m * (phk1 - phk));
                real tA_l = coef_l * (eps_p * (rhs[f] - rhs[f - 1]) + eps_m * (phk - phkm));
                real cqw_f = W->moist ? cq1_of(W->moist->cqwr[IM(f, j, i)]) : cqw_at(f, nz);
                real termA = ((msft_inv * cqw_f) * (c05dtg * g->rdn[f])) * (tA_u - tA_l);
                real bu = (W->c2a[IM(f, j, i)] * W->alt[IM(f, j, i)]) * t2n[f];
                real bl = (W->c2a[IM(f - 1, j, i)] * W->alt[IM(f - 1, j, i)]) * t2n[f - 1];
                real termB = ((dts * G) * msft_inv) * (g->rdn[f] * (bu - bl) - (g->c1f[f] * muave));
                wn[f] = wn[f] + (termA + termB);
            }
            /* top face: top_lid -> 0 */
            wn[nz] = 0.0;
            wn[0] = w_surface;
            /* Thomas */
            for (int f = 1; f <= nz; ++f)
                wn[f] = (wn[f] - W->ca[IM(f, j, i)] * wn[f - 1]) * W->calpha[IM(f, j, i)];
            for (int f = nz - 1; f >= 1; --f) wn[f] = wn[f] - W->cgamma[IM(f, j, i)] * wn[f + 1];
            /* damp_opt = 3 */
            if (nl->damp_opt == 3 && nl->dampcoef > 0.0) {
                real dampmag = dts * nl->dampcoef, hdepth = nl->zdamp;
                real htop = PH1T(nz) / G;
                real hbot = htop - hdepth;
                for (int f = 1; f <= nz; ++f) {
                    real hk = PH1T(f) / G;
                    real s = wd_sin(((0.5 * PI_D) * (hk - hbot)) / hdepth);
                    real ramp = s * s;
                    real dampwt = hk >= hbot ? dampmag * ramp : 0.0;
                    real mass_f = g->c1f[f] * mut + g->c2f[f];
                    wn[f] = (wn[f] - (dampwt * mass_f) * S->w[IM(f, j, i)]) / (1.0 + dampwt);
                }
            }
            /* geopotential finish */
            for (int f = 1; f <= nz; ++f) {
                real mfs = g->c1f[f] * muts + g->c2f[f];
                W->ph_work[IM(f, j, i)] = rhs[f] + (((((msfty * 0.5) * dts) * G) * eps_p) * wn[f]) / mfs;
            }
            for (int f = 0; f <= nz; ++f) W->w_work[IM(f, j, i)] = wn[f];
            for (int k = 0; k < nz; ++k) W->t2ave[IM(k, j, i)] = t2n[k];
#undef WQ
#undef PH1T
        }
        free(rhs);
        free(wn);
        free(t2n);
    }
}

static void substep_p(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    (void)nl;

    /* ---- calc_p_rho_step (smdiv), core/calc_p_rho.py:142-183 ---- */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real mu_work = W->muts[c] - W->mut[c];
                real mass_h = g->c1h[k] * W->muts[c] + g->c2h[k];
                real safe_mass = safe_pos(mass_h, 1.0e-12);
                real mu_term = g->c1h[k] * mu_work;
                real al = -(W->alt[q] * mu_term +
                              g->rdnw[k] * (W->ph_work[IM(k + 1, j, i)] - W->ph_work[q])) /
                            safe_mass;
                real theta_1 = S->theta[q] - T0;
                real ttr = T0 + theta_1;
                real safe_t = safe_pos(ttr, 1.0e-6);
                real p = W->c2a[q] *
                           (((W->alt[q] * (W->th_work[q] - mu_term * theta_1)) / (safe_mass * safe_t)) - al);
                real ptmp = p;
                p = p + SMDIV * (p - W->pm1[q]);
                W->al[q] = al;
                W->p[q] = p;
                W->pm1[q] = ptmp;
            }
}

static void stage_finish(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    (void)nl;

    /* ---- small_step_finish_wrf, core/small_step_finish.py:22-101 ---- */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                size_t q = IU(k, j, i);
                real ms = g->c1h[k] * W->muus[I2U(j, i)] + g->c2h[k];
                real mc = g->c1h[k] * W->muu[I2U(j, i)] + g->c2h[k];
                real den = fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12);
                S->u[q] = (g->msfuy[I2U(j, i)] * W->u_work[q] + S->u[q] * mc) / den;
            }
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t q = IV(k, j, i);
                real ms = g->c1h[k] * W->muvs[I2V(j, i)] + g->c2h[k];
                real mc = g->c1h[k] * W->muv[I2V(j, i)] + g->c2h[k];
                real den = fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12);
                S->v[q] = (g->msfvx[I2V(j, i)] * W->v_work[q] + S->v[q] * mc) / den;
            }
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real ms = g->c1f[k] * W->muts[c] + g->c2f[k];
                real mc = g->c1f[k] * W->mut[c] + g->c2f[k];
                real den = fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12);
                S->w[q] = (g->msfty[c] * W->w_work[q] + S->w[q] * mc) / den;
                real ph_pert = W->ph_work[q] + S->ph_pert[q];
                real ph_base = S->ph_total[q] - S->ph_pert[q];
                S->ph_total[q] = ph_base + ph_pert;
                S->ph_pert[q] = ph_pert;
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = I2(j, i), q = IM(k, j, i);
                real ms = g->c1h[k] * W->muts[c] + g->c2h[k];
                real mc = g->c1h[k] * W->mut[c] + g->c2h[k];
                real den = fabs(ms) > 1.0e-12 ? ms : (ms >= 0.0 ? 1.0e-12 : -1.0e-12);
                real t_save = S->theta[q] - T0;
                real theta_pert = (W->th_work[q] + t_save * mc) / den;
                S->theta[q] = theta_pert + T0;
                /* p from the small step, then refreshed below */
                real p_base = S->p_total[q] - S->p_pert[q];
                S->p_total[q] = p_base + W->p[q];
                S->p_pert[q] = W->p[q];
            }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real mu_base = S->mu_total[c] - S->mu_pert[c];
            S->mu_pert[c] = W->mu[c];
            S->mu_total[c] = mu_base + W->mu[c];
        }

    /* ---- _refresh_grid_p_from_finished, operational_mode.py:2169-2231,
     *      diagnose_pressure_al_alt base branch (acoustic_wrf.py:360-438) ---- */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = I2(j, i);
            real mub = W->mub[c];
            real muts = mub + S->mu_pert[c];
            for (int k = 0; k < nz; ++k) {
                size_t q = IM(k, j, i), q1 = IM(k + 1, j, i);
                real phb_k = S->ph_total[q] - S->ph_pert[q];
                real phb_k1 = S->ph_total[q1] - S->ph_pert[q1];
                real mass_h_base = g->c1h[k] * mub + g->c2h[k];
                real denom_alb = g->dnw[k] * mass_h_base;
                real safe_da = safe_pos(denom_alb, 1.0e-12);
                real alb = -(phb_k1 - phb_k) / safe_da;
                real al, alt;
                if (nl->hypsometric_opt == 2) {
                    real dph_tot = (phb_k1 + S->ph_pert[q1]) - (phb_k + S->ph_pert[q]);
                    real alb_used = log_alpha(g, k, phb_k1 - phb_k, mub);
                    al = log_alpha(g, k, dph_tot, muts) - alb_used;
                    alt = al + alb_used;
                } else {
                    real mass_w = g->c1h[k] * muts + g->c2h[k];
                    real safe_mass = safe_pos(mass_w, 1.0e-12);
                    real mu_term = (alb * g->c1h[k]) * S->mu_pert[c];
                    real geo_term = g->rdnw[k] * (S->ph_pert[q1] - S->ph_pert[q]);
                    al = -(mu_term + geo_term) / safe_mass;
                    alt = al + alb;
                }
                real ptot = p_from_theta_alt(S->theta[q], alt);
                real p_pert = ptot - W->pb[q];
                real p_base = S->p_total[q] - S->p_pert[q];
                S->p_total[q] = p_base + p_pert;
                S->p_pert[q] = p_pert;
            }
        }
}

/* The moisture's large-step update after the acoustic loop
 * (_apply_moisture_large_step, operational_mode.py:3273-3316):
 * q = (m_origin q_origin + dt_rk qt) / m_new with the column masses
 * c1h mu + c2h of the step origin and of the finished stage. */
static void moist_update(const orc_grid *g, const rstate *S, ws_t *W, double dt_rk) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    moist_t *m = W->moist;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t c = I2(j, i), q = IM(k, j, i);
                real mass_old = g->c1h[k] * W->mut0[c] + g->c2h[k];
                real mass_new = g->c1h[k] * S->mu_total[c] + g->c2h[k];
                real inv_mass_new = 1.0 / mass_new;
                for (int sp = 0; sp < N_MOIST; ++sp)
                    m->q[sp][q] = (mass_old * m->q0[sp][q] + dt_rk * m->qt[sp][q]) * inv_mass_new;
            }
}

static void rk_stage(const orc_grid *g, const orc_namelist *nl, rstate *S, ws_t *W, int rk_step,
                     double dts, int nsub) {
    stage_setup(g, nl, S, W, rk_step, dts);
    for (int it = 0; it < nsub; ++it) {
        substep_uv(g, nl, S, W, dts);
        substep_mu_theta(g, nl, S, W, dts);
        substep_w(g, nl, S, W, dts);
        substep_p(g, nl, S, W);
    }
    stage_finish(g, nl, S, W);
    /* the stage's dt_rk: dt/3, dt/2, dt (operational_mode.py:3527-3529) */
    if (W->moist && nl->moist_adv_opt != 0)
        moist_update(g, S, W, rk_step == 1 ? nl->dt / 3.0 : (rk_step == 2 ? 0.5 * nl->dt : nl->dt));
}

/* The rk1 reference of a large step: copies of the state it starts from
 * (operational_mode.py:3328-3329). */
static void step_origin(const orc_grid *g, const orc_namelist *nl, rstate *s, ws_t *W) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t M = (size_t)nz * ny * nx, Wn = (size_t)(nz + 1) * ny * nx;
    size_t U = (size_t)nz * ny * (nx + 1), V = (size_t)nz * (ny + 1) * nx, C = (size_t)ny * nx;
    (void)nl;
    memcpy(W->u0, s->u, U * sizeof(real));
    memcpy(W->v0, s->v, V * sizeof(real));
    memcpy(W->w0, s->w, Wn * sizeof(real));
    memcpy(W->th0, s->theta, M * sizeof(real));
    memcpy(W->ph0, s->ph_pert, Wn * sizeof(real));
    memcpy(W->mu0, s->mu_pert, C * sizeof(real));
    if (W->mut0) memcpy(W->mut0, s->mu_total, C * sizeof(real));
    if (W->moist)
        for (int sp = 0; sp < N_MOIST; ++sp) memcpy(W->moist->q0[sp], W->moist->q[sp], M * sizeof(real));
}

/* One large step in place: the origin, then the solve_em.F cadence of
 * substeps 1, ns/2, ns (operational_mode.py:3524-3534). */
static void large_step(const orc_grid *g, const orc_namelist *nl, rstate *s, ws_t *W) {
    const double dt = nl->dt;
    const int ns = nl->ns;
    step_origin(g, nl, s, W);
    rk_stage(g, nl, s, W, 1, dt / 3.0, 1);
    rk_stage(g, nl, s, W, 2, dt / (double)ns, ns / 2 > 1 ? ns / 2 : 1);
    rk_stage(g, nl, s, W, 3, dt / (double)ns, ns);
}

/* One piece of a large step (ORC_PIECE_* of wrfdyn_oracle.h) on S and W. */
static int run_piece(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
                     rstate *S, ws_t *W) {
    switch (piece) {
        case ORC_PIECE_ORIGIN: step_origin(g, nl, S, W); return 0;
        case ORC_PIECE_SETUP: stage_setup(g, nl, S, W, rk_step, dts); return 0;
        case ORC_PIECE_UV: substep_uv(g, nl, S, W, dts); return 0;
        case ORC_PIECE_MU_THETA: substep_mu_theta(g, nl, S, W, dts); return 0;
        case ORC_PIECE_W: substep_w(g, nl, S, W, dts); return 0;
        case ORC_PIECE_P: substep_p(g, nl, S, W); return 0;
        case ORC_PIECE_FINISH: stage_finish(g, nl, S, W); return 0;
        default: return -2;
    }
}

/* S and W over caller arrays in the order of array_names. */
static void bind_arrays(real *const *arrays, rstate *S, ws_t *W) {
    real **slot[N_ARRAYS];
    memset(W, 0, sizeof(*W));
    array_slots(S, W, slot);
    for (int a = 0; a < N_ARRAYS; ++a) *slot[a] = arrays[a];
}

int orc_array_count(void) { return N_ARRAYS; }
const char *orc_array_name(int a) { return a >= 0 && a < N_ARRAYS ? array_names[a] : ""; }
int orc_array_extent(int a) { return a >= 0 && a < N_ARRAYS ? array_extent[a] : -1; }

#ifndef ORC_AD
int orc_run(const orc_grid *g, const orc_namelist *nl, orc_state *s, int n_steps) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (nx < 4 || ny < 1 || nz < 4) return -2;
    if (!nl->top_lid) return -3;
    if (nl->hypsometric_opt == 2 && (!g->c3h || !g->c4h || !g->c3f || !g->c4f)) return -5;
    ws_t W;
    if (ws_alloc(&W, nx, ny, nz) != 0) {
        ws_free(&W);
        return -1;
    }
    rstate rs = {s->u, s->v, s->w, s->theta, s->p_total, s->p_pert,
                 s->ph_total, s->ph_pert, s->mu_total, s->mu_pert, s->mudf};
    /* the moisture: all six species or none */
    double *qs[N_MOIST] = {s->qv, s->qc, s->qr, s->qi, s->qs, s->qg};
    int nq = 0;
    for (int sp = 0; sp < N_MOIST; ++sp) nq += qs[sp] != NULL;
    if (nq != 0 && nq != N_MOIST) {
        ws_free(&W);
        return -6;
    }
    moist_t moist;
    memset(&moist, 0, sizeof moist);
    int rc = 0;
    if (nq == N_MOIST) {
        const size_t Mn = (size_t)nz * ny * nx;
        for (int sp = 0; sp < N_MOIST; ++sp) {
            moist.q[sp] = qs[sp];
            moist.q0[sp] = zalloc(Mn);
            moist.qt[sp] = zalloc(Mn);
            if (!moist.q0[sp] || !moist.qt[sp]) rc = -1;
        }
        moist.qtot = zalloc(Mn);
        moist.cqu = zalloc((size_t)nz * ny * (nx + 1));
        moist.cqv = zalloc((size_t)nz * (ny + 1) * nx);
        moist.cqwr = zalloc((size_t)(nz + 1) * ny * nx);
        if (!moist.qtot || !moist.cqu || !moist.cqv || !moist.cqwr) rc = -1;
        W.moist = &moist;
    }
    if (rc == 0)
        for (int step = 0; step < n_steps; ++step) large_step(g, nl, &rs, &W);
    if (nq == N_MOIST) {
        for (int sp = 0; sp < N_MOIST; ++sp) {
            free(moist.q0[sp]);
            free(moist.qt[sp]);
        }
        free(moist.qtot);
        free(moist.cqu);
        free(moist.cqv);
        free(moist.cqwr);
    }
    ws_free(&W);
    return rc;
}

int orc_piece(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
              double *const *arrays) {
    if (g->nx < 4 || g->ny < 1 || g->nz < 4) return -2;
    if (nl->scalar_adv_opt != 0) return -4; /* needs the step-origin mu_total */
    rstate S;
    ws_t W;
    bind_arrays(arrays, &S, &W);
    return run_piece(g, nl, piece, rk_step, dts, &S, &W);
}

/* An acoustic piece that reads the moisture's coupling factors (advance_uv:
 * cqu, cqv; advance_w: the raw cqw) run with the given factor arrays. */
int orc_piece_moist(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
                    double *const *arrays, const double *cqu, const double *cqv, const double *cqwr) {
    if (g->nx < 4 || g->ny < 1 || g->nz < 4) return -2;
    if (piece != ORC_PIECE_UV && piece != ORC_PIECE_W) return -2;
    if (!cqu || !cqv || !cqwr) return -2;
    rstate S;
    ws_t W;
    moist_t m;
    memset(&m, 0, sizeof m);
    m.cqu = (real *)cqu;
    m.cqv = (real *)cqv;
    m.cqwr = (real *)cqwr;
    bind_arrays(arrays, &S, &W);
    W.moist = &m;
    return run_piece(g, nl, piece, rk_step, dts, &S, &W);
}
#endif
