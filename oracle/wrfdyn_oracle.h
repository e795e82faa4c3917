/*
 * CPU oracle for the split-explicit ARW dycore large step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is never linked into, loaded by or
 * called from the product (wrf_gpu_amd/, include/).  Only tests/, the
 * smoke() check in __graft_entry__.py and bench.py's cpu_baseline leg use it.
 *
 * It restates, function by function and in the same expression order, the
 * reference's dry periodic idealized path (physics and lateral boundaries
 * off, hypsometric_opt=1, unit or general map factors, f=e=0, q=0):
 *   src/gpuwrf/runtime/operational_mode.py:3319-3534  _rk_scan_step
 * Each function in wrfdyn_oracle.c cites the reference file:line it follows.
 *
 * Layout: (k, j, i) C order, i contiguous, exactly the reference's
 *   u (nz, ny, nx+1), v (nz, ny+1, nx), w/ph (nz+1, ny, nx),
 *   mass fields (nz, ny, nx), column fields (ny, nx).
 */
#ifndef WRFDYN_ORACLE_H
#define WRFDYN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    double dx, dy;
    /* vertical metrics (contracts/grid.py:166-245) */
    const double *c1h, *c2h;                 /* nz   */
    const double *c1f, *c2f;                 /* nz+1 */
    const double *dnw, *rdnw, *dn, *rdn;     /* nz   */
    const double *fnm, *fnp;                 /* nz   */
    double cf1, cf2, cf3;
    /* map factors */
    const double *msftx, *msfty;             /* ny*nx      */
    const double *msfux, *msfuy;             /* ny*(nx+1)  */
    const double *msfvx, *msfvy;             /* (ny+1)*nx  */
    /* dry-pressure coefficients p = c3 mu + c4 + p_top, read only with
     * hypsometric_opt 2 (may be NULL otherwise) */
    const double *c3h, *c4h;                 /* nz   */
    const double *c3f, *c4f;                 /* nz+1 */
    double p_top;
    /* Coriolis parameters on mass points (DycoreMetrics f, e, sina, cosa,
     * contracts/grid.py:65-73): f = 2 Omega sin(lat), e = 2 Omega cos(lat)
     * and the projection rotation.  NULL (all four) means no rotation:
     * f = e = sina = 0, cosa = 1, and every Coriolis term is skipped. */
    const double *f, *e, *sina, *cosa;       /* ny*nx      */
} orc_grid;

typedef struct {
    double dt;          /* large step dt_s                          */
    int    ns;          /* namelist acoustic_substeps               */
    double epssm;
    int    top_lid;     /* only 1 is supported                      */
    int    w_damping;   /* 0 or 1                                   */
    int    damp_opt;    /* 0 or 3                                   */
    double dampcoef;
    double zdamp;
    int    diff_6th_opt;     /* 0 off; 1 or 2: 6th-order filter (monotonic) */
    double diff_6th_factor;
    int    diff_opt;    /* 0 off; 1 with km_opt 4: 2-D Smagorinsky horizontal diffusion */
    int    km_opt;      /* 0 or 4                                   */
    double c_s;         /* Smagorinsky constant (reference default 0.25) */
    int    scalar_adv_opt; /* theta advection on the final RK stage: 0 plain,
                              1 positive-definite, 2 monotonic limiter */
    int    hypsometric_opt; /* 1 linear al (dp = mu d eta), 2 log-pressure al */
    double const_nu_m2_s;   /* > 0: constant-K flux-form diffusion, K in m^2/s */
    int    moist_adv_opt;   /* moisture transport: 0 off, 1 positive-definite,
                               2 monotonic limiter on the final stage (plain
                               h5/v3 on the others) */
} orc_namelist;

typedef struct {
    double *u, *v, *w, *theta;
    double *p_total, *p_pert;
    double *ph_total, *ph_pert;
    double *mu_total, *mu_pert;
    double *mudf;
    /* the six moisture species (mixing ratios, nz*ny*nx each), all six or
     * none: NULL is the dry state */
    double *qv, *qc, *qr, *qi, *qs, *qg;
} orc_state;

/* Advance n_steps large steps in place.  Returns 0 or a negative error. */
int orc_run(const orc_grid *g, const orc_namelist *nl, orc_state *s, int n_steps);

/* The arrays a large step computes on -- the 11 state fields (orc_state's
 * order), then the workspace -- by index: name, and extent code
 * 0 (nz,ny,nx), 1 (nz+1,ny,nx), 2 (nz,ny,nx+1), 3 (nz,ny+1,nx), 4 (ny,nx),
 * 5 (ny,nx+1), 6 (ny+1,nx). */
int orc_array_count(void);
const char *orc_array_name(int a);
int orc_array_extent(int a);

/* A large step is ORC_PIECE_ORIGIN, then per RK stage (rk_step 1..3, dts
 * dt/3, dt/ns, dt/ns) ORC_PIECE_SETUP, nsub (1, max(ns/2,1), ns) times the
 * four substep pieces UV, MU_THETA, W, P, and ORC_PIECE_FINISH.  orc_piece
 * runs one piece in place on `arrays` (orc_array_count() pointers in the
 * order above).  Returns 0 or a negative error. */
#define ORC_PIECE_ORIGIN 0
#define ORC_PIECE_SETUP 1
#define ORC_PIECE_UV 2
#define ORC_PIECE_MU_THETA 3
#define ORC_PIECE_W 4
#define ORC_PIECE_P 5
#define ORC_PIECE_FINISH 6
int orc_piece(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
              double *const *arrays);
/* ORC_PIECE_UV or ORC_PIECE_W with the moisture's coupling factors: cqu
 * (nz, ny, nx+1), cqv (nz, ny+1, nx) and the raw cqw (nz+1, ny, nx), as
 * orc_moist_factors gives them; for the pins of those two uses. */
int orc_piece_moist(const orc_grid *g, const orc_namelist *nl, int piece, int rk_step, double dts,
                    double *const *arrays, const double *cqu, const double *cqv, const double *cqwr);

/* Single-operator entry points used by the oracle's own pinning tests. */
int orc_calc_coef_w(const orc_grid *g, const double *mut, const double *c2a,
                    double dts, double epssm, int top_lid,
                    double *a, double *alpha, double *gamma);
/* the same with the moisture: cqw_raw (nz+1, ny, nx) the raw calc_cq face
 * field (0.5 (qtot(k) + qtot(k-1)) inside), NULL for the dry state */
int orc_calc_coef_w_cq(const orc_grid *g, const double *mut, const double *c2a, double dts, double epssm,
                       int top_lid, const double *cqw_raw, double *a, double *alpha, double *gamma);
void orc_thomas(int nzp1, const double *a, const double *alpha, const double *gamma,
                const double *rhs, double *out);
double orc_alt_from_theta_p(double theta, double p);
double orc_p_from_theta_alt(double theta, double alt);
int orc_advect_scalar(const orc_grid *g, const double *u, const double *v,
                      const double *mu_total, const double *field, double *tend);
/* Terrain surface w (advance_w.py:505-556) for every column: phb0 is the
 * surface base geopotential (ny*nx), u and v the staggered 3-D winds. */
int orc_surface_w(const orc_grid *g, const double *phb0, const double *u, const double *v, double *out);
/* The uncoupled 6th-order filter tendency (explicit_diffusion.py:200-223)
 * of a (nlev, nrow, ncol) field, periodic in rows and columns. */
int orc_sixth_tendency(int nlev, int nrow, int ncol, const double *f, double coef, double *out);
/* The coupled tendencies of the 2-D Smagorinsky horizontal diffusion
 * (diff_opt 1, km_opt 4; operational_mode.py:2946-2963) of state s alone,
 * with its eddy viscosity xkmh (nz*ny*nx) when xkmh is not NULL. */
/* The coupled tendency of q with the positive-definite (opt 1) or monotonic
 * (opt 2) limiter (flux_advection.py:736-975), transport from u, v and
 * mu_total, limits from the origin field qold and mass mu_old over dt. */
int orc_advect_limited(const orc_grid *g, int opt, const double *u, const double *v, const double *mu_total,
                       const double *q, const double *qold, const double *mu_old, double dt, double *tend);
/* The constant-K diffusion tendencies alone (zero + the diffusion) of
 * state s with K = nu, and the layer thickness dz it uses; 0 on success. */
int orc_constk_tendency(const orc_grid *g, double nu, const orc_state *s, double *ut, double *vt, double *wt,
                        double *tht, double *dz);
int orc_moist_factors(const orc_grid *g, const orc_state *s, double *qtot, double *cqu, double *cqv, double *cqwr);
int orc_coriolis_tendency(const orc_grid *g, const orc_state *s, double *ut, double *vt, double *rwt);
int orc_smag_tendency(const orc_grid *g, double c_s, const orc_state *s, double *ut, double *vt, double *wt,
                      double *tht, double *xkmh);
double orc_pow(double x, double y);
double orc_sin(double x);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
