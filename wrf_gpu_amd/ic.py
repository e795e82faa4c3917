"""Idealized initial conditions for the dycore workloads (FP64, seeded).

The column construction follows the reference's idealized generator
(src/gpuwrf/ic_generators/idealized.py):

* ``uniform_z_hydrostatic_base`` restates ``_uniform_z_hydrostatic_base``
  (:72-110) for the 2-D Skamarock warm bubble slab (:219-274).
* ``_balanced_state`` restates ``_make_state`` (:468-563) -- base geopotential
  from the base inverse density, fixed-mass hydrostatic re-integration of the
  perturbation geopotential after the theta perturbation, and the start_em
  equivalent perturbation pressure from the EOS -- generalised from one slab
  column profile to a column per (j, i), so it also covers terrain (a
  per-column dry mass) and a stratified base state.

The IC builder uses the reference generator's constants (R=287, cp=1004),
which differ from the dycore's cp=7R/2 on purpose (idealized.py:53-58).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import Metrics, Namelist

GRAVITY = 9.81
R_DRY = 287.0
CP_DRY = 1004.0
CV_DRY = CP_DRY - R_DRY
P0 = 100000.0
THETA0 = 300.0


@dataclass
class Case:
    """A dycore workload: metrics, namelist and the initial state dict."""

    name: str
    metrics: Metrics
    namelist: Namelist
    state: dict
    n_steps: int
    info: dict


def alpha_dry(theta, pressure, cp=CP_DRY):
    """``_alpha_dry``, idealized.py:61-69: alt = (R/p0)*theta*(p/p0)^(-cv/cp)."""

    theta = np.asarray(theta, dtype=np.float64)
    pressure = np.asarray(pressure, dtype=np.float64)
    cv = cp - R_DRY
    return (R_DRY / P0) * theta * (pressure / P0) ** (-cv / cp)


def centered_cosine(radius):
    """``_centered_cosine``, idealized.py:201-203."""

    radius = np.asarray(radius, dtype=np.float64)
    return np.where(radius <= 1.0, 0.5 * (1.0 + np.cos(np.pi * radius)), 0.0)


def uniform_z_hydrostatic_base(z_face, theta0, p_surface=P0):
    """``_uniform_z_hydrostatic_base``, idealized.py:72-110."""

    z_face = np.asarray(z_face, dtype=np.float64)
    nz = int(z_face.size) - 1
    p_face = np.zeros(nz + 1, dtype=np.float64)
    p_face[0] = float(p_surface)
    for k in range(nz):
        dz = z_face[k + 1] - z_face[k]
        alpha_lo = alpha_dry(np.array([theta0]), np.array([p_face[k]]))[0]
        p_mid = p_face[k] - 0.5 * GRAVITY * dz / alpha_lo
        alpha_mid = alpha_dry(np.array([theta0]), np.array([p_mid]))[0]
        p_face[k + 1] = p_face[k] - GRAVITY * dz / alpha_mid
    p_top = float(p_face[-1])
    mu = float(p_face[0] - p_top)
    eta = (p_face - p_top) / mu
    ph_face = GRAVITY * z_face
    p_mass = 0.5 * (p_face[:-1] + p_face[1:])
    return eta, p_mass, ph_face, mu


def _balanced_state(metrics: Metrics, *, mub, pb, alb, theta_full, phb_surface, u0=0.0, v0=0.0,
                    cp=CP_DRY):
    """Per-column restatement of ``_make_state`` (idealized.py:468-563).

    ``mub`` (ny, nx) base dry mass, ``pb`` (nz, ny, nx) base pressure at mass
    levels, ``alb`` (nz, ny, nx) base inverse density, ``theta_full``
    (nz, ny, nx) total potential temperature, ``phb_surface`` (ny, nx)
    surface geopotential.  Pure sigma (c1h=1, c2h=0) and mu' = 0.
    """

    nz, ny, nx = metrics.nz, metrics.ny, metrics.nx
    wrf_dnw = metrics.eta[1:] - metrics.eta[:-1]
    mass_h = np.asarray(mub, dtype=np.float64)
    alt_full = alpha_dry(theta_full, pb, cp)
    al = alt_full - alb
    phb = np.zeros((nz + 1, ny, nx), dtype=np.float64)
    ph_pert = np.zeros((nz + 1, ny, nx), dtype=np.float64)
    phb[0] = phb_surface
    for k in range(nz):
        phb[k + 1] = phb[k] - wrf_dnw[k] * mass_h * alb[k]
        ph_pert[k + 1] = ph_pert[k] - wrf_dnw[k] * mass_h * al[k]
    wrf_rdnw = 1.0 / wrf_dnw
    al_diag = -(wrf_rdnw[:, None, None] * (ph_pert[1:] - ph_pert[:-1])) / mass_h
    alt_diag = al_diag + alb
    cpovcv = cp / (cp - R_DRY)
    p_full = P0 * ((R_DRY * theta_full) / (P0 * alt_diag)) ** cpovcv
    p_pert = p_full - pb
    state = {
        "u": np.full((nz, ny, nx + 1), float(u0)),
        "v": np.full((nz, ny + 1, nx), float(v0)),
        "w": np.zeros((nz + 1, ny, nx)),
        "theta": np.ascontiguousarray(theta_full, dtype=np.float64),
        "p_total": pb + p_pert,
        "p_perturbation": p_pert,
        "ph_total": phb + ph_pert,
        "ph_perturbation": ph_pert,
        "mu_total": np.array(mass_h, dtype=np.float64, copy=True),
        "mu_perturbation": np.zeros((ny, nx)),
        "mudf": np.zeros((ny, nx)),
    }
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in state.items()}


def warm_bubble_2d() -> Case:
    """The reference's 2-D Skamarock warm bubble slab (idealized.py:219-274, :566-602).

    80 x 40 (ny = 1), dx = dz = 250 m, dt = 0.1 s, 10 acoustic substeps,
    rigid lid, w_damping=1, damp_opt=3, dampcoef=0.2, zdamp=3000 m.
    """

    dx = 250.0
    nx, nz = 80, 40
    x = (np.arange(nx) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    z_face = np.arange(nz + 1) * dx
    xx, zz = np.meshgrid(x, z)
    radius = np.sqrt(((xx - 10000.0) / 2000.0) ** 2 + ((zz - 2000.0) / 2000.0) ** 2)
    theta_prime = 2.0 * centered_cosine(radius)
    theta_prime *= 2.0 / float(np.max(theta_prime))
    eta, p_mass, _ph_face, mu = uniform_z_hydrostatic_base(z_face, THETA0)
    p_top = float(P0 - mu)
    metrics = Metrics.flat(nx=nx, ny=1, nz=nz, dx=dx, dy=dx, eta=eta, p_top=p_top)
    theta = (THETA0 + theta_prime)[:, None, :]
    pb = np.broadcast_to(p_mass[:, None, None], (nz, 1, nx)).copy()
    alb = np.broadcast_to(alpha_dry(np.full(nz, THETA0), p_mass)[:, None, None], (nz, 1, nx)).copy()
    state = _balanced_state(metrics, mub=np.full((1, nx), mu), pb=pb, alb=alb, theta_full=theta,
                            phb_surface=np.zeros((1, nx)))
    nl = Namelist(dt_s=0.1, acoustic_substeps=10, zdamp=3000.0)
    return Case("warm_bubble_2d", metrics, nl, state, 5000,
                {"x_m": x, "z_m": z, "snapshot_steps": (1000, 2500, 5000)})


def with_moisture(case: Case, seed: int = 7, qv_surface: float = 0.012, moist_adv_opt: int = 1) -> Case:
    """``case`` with the six moisture species added (a moist state): qv decaying with height
    from ``qv_surface`` (2.5 km scale) with seeded 5 % noise, a cloud-water blob of 1 g/kg and
    small rain, ice, snow and graupel blobs beside it, all non-negative, and the namelist's
    ``moist`` flag and ``moist_adv_opt`` set."""

    m = case.metrics
    nz, ny, nx = m.nz, m.ny, m.nx
    ph = case.state["ph_total"]
    z = 0.5 * (ph[:-1] + ph[1:]) / GRAVITY
    rng = np.random.default_rng(seed)
    qv = qv_surface * np.exp(-np.maximum(z - z[0:1], 0.0) / 2500.0) * (1.0 + 0.05 * rng.uniform(-1, 1, (nz, ny, nx)))
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    kk = np.arange(nz)[:, None, None]

    def blob(ci, cj, ck, r, amp):
        d = np.sqrt(((ii[None] - ci) / r) ** 2 + ((jj[None] - cj) / max(r * ny / nx, 0.5)) ** 2 + ((kk - ck) / 2.0) ** 2)
        return amp * np.where(d < 1.0, np.cos(0.5 * np.pi * d) ** 2, 0.0)

    state = dict(case.state)
    state["qv"] = np.ascontiguousarray(qv)
    state["qc"] = np.ascontiguousarray(blob(nx * 0.4, ny * 0.5, nz * 0.3, max(nx / 6, 1.5), 1.0e-3))
    state["qr"] = np.ascontiguousarray(blob(nx * 0.6, ny * 0.4, nz * 0.2, max(nx / 8, 1.5), 4.0e-4))
    state["qi"] = np.ascontiguousarray(blob(nx * 0.5, ny * 0.6, nz * 0.7, max(nx / 7, 1.5), 2.0e-4))
    state["qs"] = np.ascontiguousarray(blob(nx * 0.3, ny * 0.3, nz * 0.6, max(nx / 9, 1.5), 1.0e-4))
    state["qg"] = np.ascontiguousarray(blob(nx * 0.7, ny * 0.7, nz * 0.4, max(nx / 9, 1.5), 1.5e-4))
    nl = case.namelist.replace(moist=True, moist_adv_opt=int(moist_adv_opt))
    return Case(case.name + "_moist", m, nl, state, case.n_steps, dict(case.info, moist_seed=seed))


def density_current_2d() -> Case:
    """The reference's 2-D Straka density current slab (idealized.py:277-330, :566-602).

    500 x 60 (ny = 1), dx = dz = 100 m, x from -25 km; a -15 K cos^2 cold bubble of
    radii 4 km (x) and 2 km (z) centred at (0, 3 km) on the neutral theta0 = 300 K
    base; dt = 0.1 s, 10 acoustic substeps, rigid lid, w_damping=1, damp_opt=3,
    dampcoef=0.2, zdamp=3000 m, constant-K diffusion with K = 75 m^2/s.
    """

    dx = 100.0
    nx, nz = 500, 60
    x = -25000.0 + (np.arange(nx) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    z_face = np.arange(nz + 1) * dx
    xx, zz = np.meshgrid(x, z)
    radius = np.sqrt((xx / 4000.0) ** 2 + ((zz - 3000.0) / 2000.0) ** 2)
    theta_prime = -15.0 * centered_cosine(radius)
    theta_prime *= 15.0 / abs(float(np.min(theta_prime)))
    eta, p_mass, _ph_face, mu = uniform_z_hydrostatic_base(z_face, THETA0)
    p_top = float(P0 - mu)
    metrics = Metrics.flat(nx=nx, ny=1, nz=nz, dx=dx, dy=dx, eta=eta, p_top=p_top)
    theta = (THETA0 + theta_prime)[:, None, :]
    pb = np.broadcast_to(p_mass[:, None, None], (nz, 1, nx)).copy()
    alb = np.broadcast_to(alpha_dry(np.full(nz, THETA0), p_mass)[:, None, None], (nz, 1, nx)).copy()
    state = _balanced_state(metrics, mub=np.full((1, nx), mu), pb=pb, alb=alb, theta_full=theta,
                            phb_surface=np.zeros((1, nx)))
    nl = Namelist(dt_s=0.1, acoustic_substeps=10, zdamp=3000.0, const_nu_m2_s=75.0)
    return Case("density_current_2d", metrics, nl, state, 9000,
                {"x_m": x, "z_m": z, "snapshot_steps": (9000,)})


def _column_base(metrics: Metrics, mub, phb_surface, theta_of_z, iterations=8, cp=CP_DRY):
    """Hydrostatic base columns on the eta levels for a theta(z) profile.

    pb = eta_mass*mub + p_top (pure sigma), alb = alpha(theta(z_mass), pb),
    phb(k+1) = phb(k) - dnw(k)*mub*alb(k) (idealized.py:514-518); z_mass is
    iterated to a fixed point because theta depends on height.
    """

    nz = metrics.nz
    eta = metrics.eta
    eta_mass = 0.5 * (eta[:-1] + eta[1:])
    dnw = eta[1:] - eta[:-1]
    pb = eta_mass[:, None, None] * mub[None] + metrics.p_top
    # first guess: isothermal-ish heights from the pressure
    z_mass = (phb_surface / GRAVITY)[None] + 7000.0 * np.log(pb[0:1] / pb)
    for _ in range(int(iterations)):
        alb = alpha_dry(theta_of_z(z_mass), pb, cp)
        phb = np.empty((nz + 1,) + mub.shape)
        phb[0] = phb_surface
        for k in range(nz):
            phb[k + 1] = phb[k] - dnw[k] * mub * alb[k]
        z_mass = 0.5 * (phb[:-1] + phb[1:]) / GRAVITY
    alb = alpha_dry(theta_of_z(z_mass), pb, cp)
    return pb, alb, z_mass


def bubble_3d(nx=64, ny=64, nz=40, dx=1000.0, p_top=5000.0, u0=10.0, noise=0.01, seed=1,
              dt=6.0, ns=6, zdamp=5000.0, amplitude=2.0, cp=CP_DRY, dy=None) -> Case:
    """Plain FP64 dry 3-D warm bubble (the 64x64x40 CPU-runnable workload).

    Flat, uniform eta, neutral theta0=300 K hydrostatic from 1e5 Pa (mu=95000 Pa),
    +2 K cos^2 bubble of radius 2 km centred at (domain centre, 2 km),
    uniform u, seeded uniform +-noise K theta perturbation.
    """

    dy = dx if dy is None else dy
    eta = np.linspace(1.0, 0.0, nz + 1)
    metrics = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, eta=eta, p_top=p_top)
    mub = np.full((ny, nx), P0 - p_top)
    pb, alb, z_mass = _column_base(metrics, mub, np.zeros((ny, nx)), lambda z: np.full_like(z, THETA0),
                                   cp=cp)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    xc, yc = 0.5 * nx * dx, 0.5 * ny * dy
    r = np.sqrt(((x[None, None, :] - xc) / 2000.0) ** 2 + ((y[None, :, None] - yc) / 2000.0) ** 2
                + ((z_mass - 2000.0) / 2000.0) ** 2)
    tp = 2.0 * centered_cosine(r)
    if np.max(tp) > 0 and amplitude != 0.0:
        tp *= amplitude / float(np.max(tp))
    else:
        tp *= 0.0
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        tp = tp + rng.uniform(-noise, noise, size=(nz, ny, nx))
    theta = THETA0 + tp
    state = _balanced_state(metrics, mub=mub, pb=pb, alb=alb, theta_full=theta,
                            phb_surface=np.zeros((ny, nx)), u0=u0, cp=cp)
    nl = Namelist(dt_s=float(dt), acoustic_substeps=int(ns), zdamp=float(zdamp))
    return Case(f"bubble_{nx}x{ny}x{nz}", metrics, nl, state, 100, {"seed": seed})


def _correlated_field(shape, sigma, length_m, dx, rng):
    """Gaussian random field, horizontally correlated by an FFT Gaussian filter."""

    nz, ny, nx = shape
    white = rng.standard_normal(shape)
    kx = np.fft.fftfreq(nx, d=dx) * 2.0 * np.pi
    ky = np.fft.fftfreq(ny, d=dx) * 2.0 * np.pi
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = np.exp(-0.5 * k2 * length_m ** 2)
    out = np.real(np.fft.ifft2(np.fft.fft2(white, axes=(1, 2)) * filt[None], axes=(1, 2)))
    std = float(np.std(out))
    return out * (sigma / std) if std > 0 else out


def stratified_terrain_3d(nx=433, ny=433, nz=50, dx=3000.0, p_top=5000.0, u0=10.0, seed=2,
                          gamma=1.4, dt=18.0, ns=6, zdamp=5000.0, sigma=0.1,
                          length_m=20000.0, u_sigma=0.0) -> Case:
    """Stratified flow over a bell-shaped hill (the 433x433x50 workload).

    Stretched eta ``eta_k = 1 - (k/nz)**gamma`` (finer near the surface),
    h = 1000/(1+(r/10 km)^2) m, theta(z) = 288 exp(N^2 z/g) with N = 0.01 /s,
    uniform u, seeded Gaussian theta' (sigma, 20 km correlation).
    """

    eta = 1.0 - (np.arange(nz + 1) / nz) ** gamma
    metrics = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, eta=eta, p_top=p_top)
    n2 = 0.01 ** 2
    theta_s = 288.0

    def theta_of_z(z):
        return theta_s * np.exp(n2 * z / GRAVITY)

    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    xc, yc = 0.5 * nx * dx, 0.5 * ny * dx
    rr = np.sqrt((x[None, :] - xc) ** 2 + (y[:, None] - yc) ** 2)
    h = 1000.0 / (1.0 + (rr / 10000.0) ** 2)
    # surface pressure of the N=const atmosphere at height h (analytic Exner)
    kappa = R_DRY / CP_DRY
    exner = 1.0 - GRAVITY ** 2 / (CP_DRY * theta_s * n2) * (1.0 - np.exp(-n2 * h / GRAVITY))
    p_sfc = P0 * exner ** (1.0 / kappa)
    mub = p_sfc - p_top
    pb, alb, z_mass = _column_base(metrics, mub, GRAVITY * h, theta_of_z)
    rng = np.random.default_rng(seed)
    tp = _correlated_field((nz, ny, nx), sigma, length_m, dx, rng)
    theta = theta_of_z(z_mass) + tp
    state = _balanced_state(metrics, mub=mub, pb=pb, alb=alb, theta_full=theta,
                            phb_surface=GRAVITY * h, u0=u0)
    if u_sigma > 0.0:
        state["u"] = state["u"] + _correlated_field(state["u"].shape, u_sigma, length_m, dx, rng)
    nl = Namelist(dt_s=float(dt), acoustic_substeps=int(ns), zdamp=float(zdamp))
    return Case(f"terrain_{nx}x{ny}x{nz}", metrics, nl, state, 200,
                {"seed": seed, "eta_gamma": gamma})


def ensemble_member(nx, ny, nz=50, member=0, dx=1000.0, p_top=5000.0, u0=10.0, sigma=0.5,
                    u_sigma=1.0, length_m=20000.0, dt=6.0, ns=6, zdamp=5000.0) -> Case:
    """One member of the uneven-size ensemble: the flat, uniform-eta, neutral
    theta0=300 K base state of the warm-bubble workload with seeded, 20 km
    correlated theta' (sigma 0.5 K) and u' (sigma 1 m/s) perturbations; the
    member's seed is 100 + member."""

    seed = 100 + int(member)
    eta = np.linspace(1.0, 0.0, nz + 1)
    metrics = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, eta=eta, p_top=p_top)
    mub = np.full((ny, nx), P0 - p_top)
    pb, alb, _ = _column_base(metrics, mub, np.zeros((ny, nx)), lambda z: np.full_like(z, THETA0))
    rng = np.random.default_rng(seed)
    tp = _correlated_field((nz, ny, nx), sigma, length_m, dx, rng)
    state = _balanced_state(metrics, mub=mub, pb=pb, alb=alb, theta_full=THETA0 + tp,
                            phb_surface=np.zeros((ny, nx)), u0=u0)
    state["u"] = state["u"] + _correlated_field(state["u"].shape, u_sigma, length_m, dx, rng)
    nl = Namelist(dt_s=float(dt), acoustic_substeps=int(ns), zdamp=float(zdamp))
    return Case(f"member{member}_{nx}x{ny}x{nz}", metrics, nl, state, 200, {"seed": seed})


ENSEMBLE_SIZES = ((129, 129), (201, 201), (257, 257), (321, 321))


def ensemble_32(nz=50, per_size=8):
    """The 32-member uneven ensemble: ``per_size`` members each of 129^2,
    201^2, 257^2 and 321^2 columns, ``nz`` levels (members numbered in that order)."""

    out = []
    for nx, ny in ENSEMBLE_SIZES:
        for _ in range(per_size):
            out.append(ensemble_member(nx, ny, nz, member=len(out)))
    return out
