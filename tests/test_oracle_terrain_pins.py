"""Pins of the oracle's terrain and stretched-eta terms, which the flat, uniform-eta golden
warm bubble leaves open (the 433x433x50 workload exercises all of them).

* Large-step horizontal PGF (core/rk_addtend_dry.py:232-361) over terrain on stretched eta:
  with smooth analytic fields (a varying base geopotential and mass, non-hydrostatic p'), the
  discrete tendency converges at second order to WRF's continuous perturbation-form PGF
      -(msfux/msfuy) [phi'_x + alpha p'_x + alpha' pb_x + phi_x (p'_eta - c1 mu') / mu]
  (all x-derivatives along eta surfaces), as the grid is refined in x and eta together.
  A sign, face-weight or index mistake in any of the four terms stops the convergence.
* rhs_ph (core/rhs_ph.py:175-398) over terrain on stretched eta: for a flow whose omega
  follows in closed form from the column continuity, the phi tendency converges at second
  order to mu g w - mu u phi_x - omega phi_eta.
* The w-curvature term (operational_mode.py:1784-1822) of uniform flow is, in closed form,
  mu (U^2 + V^2) / R_earth at every interior face.
* A resting atmosphere on stretched eta whose total state (stratified) differs from its base
  state (neutral) stays at rest: the perturbation pressure's discrete hydrostatic balance with
  mu' uses the stretched rdn, so a metric mistake sets it moving.
* w_damping (core/advance_w.py:92-132): where the vertical Courant number |ww rdnw dts / m|
  exceeds 1, the explicit w gains -sign(w) 0.3 (cfl - 1) m dts, and nothing below 1.
* hypsometric_opt 2 refresh (acoustic_wrf.py:394-427): an isothermal column at rest over a
  different isothermal base gets back exactly the dry hydrostatic pressure of its own mass
  mu_b + mu', p = c3 (mu_b + mu') + c4 + p_top, because the log form is exact for T constant.
"""

import dataclasses

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic
from wrf_gpu_amd.grid import Metrics, Namelist

R, G, P0 = 287.0, 9.81, 1.0e5
CP = 3.5 * R
CVPM = -(CP - R) / CP
RERADIUS = 1.0 / 6370.0e3
L, PTOP = 40000.0, 5000.0
KX = 2 * np.pi / L

# smooth periodic analytic fields of (x, e), e the WRF eta (1 at the surface, 0 at the top)
MUB = (lambda x: 9.0e4 + 2.0e3 * np.cos(KX * x), lambda x: -2.0e3 * KX * np.sin(KX * x))
MUP = (lambda x: 400.0 * np.sin(KX * x + 0.3), lambda x: 400.0 * KX * np.cos(KX * x + 0.3))
# base geopotential over a cosine hill of 800 m, and its x and eta derivatives
PHB = (lambda x, e: G * 800.0 * np.cos(KX * x) * e + 2.0e5 * (1 - e) + 3.0e4 * (1 - e) ** 2,
       lambda x, e: -G * 800.0 * KX * np.sin(KX * x) * e,
       lambda x, e: G * 800.0 * np.cos(KX * x) - 2.0e5 - 6.0e4 * (1 - e))
PHP = (lambda x, e: 300.0 * np.sin(KX * x + 0.7) * e * (1 - e),
       lambda x, e: 300.0 * KX * np.cos(KX * x + 0.7) * e * (1 - e),
       lambda x, e: 300.0 * np.sin(KX * x + 0.7) * (1 - 2 * e))
PP = (lambda x, e: 250.0 * np.cos(KX * x - 0.4) * (e + 0.3) ** 2,
      lambda x, e: -250.0 * KX * np.sin(KX * x - 0.4) * (e + 0.3) ** 2,
      lambda x, e: 500.0 * np.cos(KX * x - 0.4) * (e + 0.3))
TH = lambda x, e: 300.0 + 20.0 * (1 - e) + 1.5 * np.sin(KX * x) * e  # noqa: E731


def _smooth_eta(nz):
    """Levels finer near the surface on a smooth map t -> (t + t^2) / 2, so that refining keeps
    neighbouring layer ratios tending to 1 (the power law t^gamma does not at the surface)."""

    t = np.arange(nz + 1) / nz
    return 1.0 - 0.5 * (t + t * t)


def _analytic_case(nx, nz, gamma=1.4, u=None, w=None, eta=None):
    dx = L / nx
    eta = 1.0 - (np.arange(nz + 1) / nz) ** gamma if eta is None else eta
    m = Metrics.flat(nx=nx, ny=1, nz=nz, dx=dx, dy=dx, eta=eta, p_top=PTOP)
    em = 0.5 * (eta[1:] + eta[:-1])
    x = (np.arange(nx) + 0.5) * dx
    mub, mup = MUB[0](x), MUP[0](x)
    phb, php = PHB[0](x[None, :], eta[:, None]), PHP[0](x[None, :], eta[:, None])
    pb = em[:, None] * mub[None, :] + PTOP
    pp = PP[0](x[None, :], em[:, None])
    ex = lambda a: np.ascontiguousarray(a[:, None, :])  # noqa: E731
    st = {"u": np.zeros((nz, 1, nx + 1)) if u is None else u, "v": np.zeros((nz, 2, nx)),
          "w": np.zeros((nz + 1, 1, nx)) if w is None else w, "theta": ex(TH(x[None, :], em[:, None])),
          "p_total": ex(pb + pp), "p_perturbation": ex(pp), "ph_total": ex(phb + php), "ph_perturbation": ex(php),
          "mu_total": (mub + mup)[None, :].copy(), "mu_perturbation": mup[None, :].copy(), "mudf": np.zeros((1, nx))}
    return m, st


def _setup_piece(m, st):
    nl = Namelist(dt_s=6.0, acoustic_substeps=6, damp_opt=0)
    arr = oracle.new_arrays(m, st)
    oracle.piece(m, nl, arr, oracle.PIECE_SETUP, 1, 2.0)
    return arr


def _pgf_target(m):
    """WRF's continuous perturbation-form PGF over the face mass, at the u faces."""

    em = 0.5 * (m.eta[1:] + m.eta[:-1])
    X, E = (np.arange(m.nx + 1) * m.dx)[None, :], em[:, None]
    mub, mup = MUB[0](X), MUP[0](X)
    mut = mub + mup
    pb = E * mub + PTOP
    alpha = (R / P0) * TH(X, E) * ((pb + PP[0](X, E)) / P0) ** CVPM
    alb = -PHB[2](X, E) / mub  # base specific volume from the base hydrostatic relation
    alpha_p = -(alb * mup + PHP[2](X, E)) / mut
    phi_x = PHB[1](X, E) + PHP[1](X, E)
    return -(PHP[1](X, E) + alpha * PP[1](X, E) + alpha_p * E * MUB[1](X) + phi_x * (PP[2](X, E) - mup) / mut)


def test_terrain_pgf_converges_to_the_continuous_form_at_second_order():
    errs = []
    for nx, nz in ((16, 12), (32, 24), (64, 48), (128, 96)):
        m, st = _analytic_case(nx, nz)
        arr = _setup_piece(m, st)
        mt = st["mu_total"][0]
        muu = np.concatenate([mt[:1], 0.5 * (mt[1:] + mt[:-1]), mt[-1:]])
        got = arr["ut"][:, 0, :] / muu[None, :]
        want = _pgf_target(m)
        assert np.max(np.abs(want)) > 0.05
        errs.append(float(np.max(np.abs(got - want)[:, 1:-1])))  # interior faces
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(orders) >= 1.9, (errs, orders)
    assert errs[-1] < 1e-5, errs


def test_rhs_ph_over_terrain_converges_at_second_order():
    """u = a(x) eta on every level: the coupled divergence d(mu a)/dx eta gives, from the column
    continuity, omega = d(mu a)/dx eta (1 - eta) / 2, so the phi tendency's limit is
    mu g w - mu u phi_x - omega phi_eta, with the terrain's base geopotential in phi."""

    errs = []
    U0 = 8.0
    a = (lambda x: U0 * (1.0 + 0.3 * np.sin(KX * x)), lambda x: U0 * 0.3 * KX * np.cos(KX * x))
    for nx, nz in ((16, 12), (32, 24), (64, 48), (128, 96)):
        dx = L / nx
        eta = _smooth_eta(nz)
        em = 0.5 * (eta[1:] + eta[:-1])
        xu = np.arange(nx + 1) * dx
        x = (np.arange(nx) + 0.5) * dx
        u = (a[0](xu)[None, :] * em[:, None])[:, None, :].copy()
        w = (0.3 * np.cos(KX * x)[None, :] * np.sin(np.pi * eta)[:, None])[:, None, :].copy()
        m, st = _analytic_case(nx, nz, u=u, w=w, eta=eta)
        arr = _setup_piece(m, st)
        X, E = x[None, :], eta[1:-1, None]
        mu, mu_x = MUB[0](X) + MUP[0](X), MUB[1](X) + MUP[1](X)
        dmua = mu_x * a[0](X) + mu * a[1](X)
        omega = dmua * E * (1.0 - E) / 2.0
        phi_x = PHB[1](X, E) + PHP[1](X, E)
        phi_e = PHB[2](X, E) + PHP[2](X, E)
        want = mu * G * w[1:-1, 0, :] - mu * a[0](X) * E * phi_x - omega * phi_e
        assert np.max(np.abs(omega * phi_e)) > 0.1 * np.max(np.abs(want))
        got = arr["ph_tend"][1:-1, 0, :]
        errs.append(float(np.max(np.abs(got - want))) / float(np.max(np.abs(want))))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(orders) >= 1.9, (errs, orders)


def test_curvature_of_uniform_flow_is_mass_times_speed_squared_over_radius():
    U, V = 12.0, -5.0
    case = ic.bubble_3d(nx=10, ny=8, nz=7, dx=2000.0, u0=U, noise=0.0, amplitude=0.0, cp=CP)
    st = {k: np.array(v, copy=True) for k, v in case.state.items()}
    st["v"][:] = V
    arr = _setup_piece(case.metrics, st)
    mu = st["mu_total"][None]
    # rw_tend = (hydrostatic balance, zero to rounding) + w advection (zero) + curvature
    want = np.broadcast_to(mu * (U * U + V * V) * RERADIUS, arr["rw_tend"][1:-1].shape)
    assert np.allclose(arr["rw_tend"][1:-1], want, rtol=1e-9, atol=1e-9 * np.max(want))
    assert np.max(np.abs(arr["rw_tend"][[0, -1]])) < 1e-9 * np.max(want) + 1e-9


def test_rest_on_stretched_eta_with_a_different_base_stays_at_rest():
    nx, ny, nz = 6, 5, 20
    eta = 1.0 - (np.arange(nz + 1) / nz) ** 1.6
    m = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=3000.0, dy=3000.0, eta=eta, p_top=PTOP)
    zero = np.zeros((ny, nx))
    ps_b, ps_t = 1.0e5, 0.985e5
    pb, alb, _ = ic._column_base(m, np.full((ny, nx), ps_b - PTOP), zero, lambda z: np.full_like(z, 300.0),
                                 iterations=40, cp=CP)
    n2, ths = 1.0e-4, 288.0
    pt, alt, zt = ic._column_base(m, np.full((ny, nx), ps_t - PTOP), zero, lambda z: ths * np.exp(n2 * z / G),
                                  iterations=40, cp=CP)
    dnw = eta[1:] - eta[:-1]

    def ph_of(mu, al):
        ph = np.zeros((nz + 1, ny, nx))
        for k in range(nz):
            ph[k + 1] = ph[k] - dnw[k] * mu * al[k]
        return ph

    phb, pht = ph_of(ps_b - PTOP, alb), ph_of(ps_t - PTOP, alt)
    theta = ths * np.exp(n2 * zt / G)
    p_eos = P0 * ((R * theta) / (P0 * alt)) ** (CP / (CP - R))
    st = {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)), "w": np.zeros((nz + 1, ny, nx)),
          "theta": theta, "p_total": p_eos, "p_perturbation": p_eos - pb, "ph_total": pht,
          "ph_perturbation": pht - phb, "mu_total": np.full((ny, nx), ps_t - PTOP),
          "mu_perturbation": np.full((ny, nx), ps_t - ps_b), "mudf": np.zeros((ny, nx))}
    assert np.max(np.abs(st["p_perturbation"])) > 500.0 and np.max(np.abs(st["ph_perturbation"])) > 100.0
    nl = Namelist(dt_s=12.0, acoustic_substeps=6, zdamp=4000.0)
    out = oracle.run(m, nl, st, 10)
    assert np.max(np.abs(out["w"])) < 1e-9
    assert np.max(np.abs(out["u"])) < 1e-9 and np.max(np.abs(out["v"])) < 1e-9
    assert np.max(np.abs(out["ph_perturbation"] - st["ph_perturbation"])) < 1e-7
    assert np.max(np.abs(out["p_perturbation"] - st["p_perturbation"])) < 1e-7
    assert np.max(np.abs(out["theta"] - st["theta"])) < 1e-9


def _traced_arrays(case, nl):
    """The step's arrays right before the first advance_w of the first stage."""

    arr = oracle.new_arrays(case.metrics, case.state)
    for pid, rk, dts in oracle.step_pieces(nl):
        if pid == oracle.PIECE_W:
            return arr, rk, dts
        oracle.piece(case.metrics, nl, arr, pid, rk, dts)
    raise AssertionError("no advance_w piece")


def test_w_damping_acts_above_a_vertical_courant_number_of_one():
    case = ic.stratified_terrain_3d(nx=8, ny=6, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = case.metrics
    base = dataclasses.replace(case.namelist, damp_opt=0)
    arr, rk, dts = _traced_arrays(case, base)
    nz = m.nz
    mut = arr["mut"]
    mass_f = m.c1f[:, None, None] * mut[None] + m.c2f[:, None, None]
    # omega giving a known vertical Courant number on every interior face, 0.5 to 2.5
    rng = np.random.default_rng(4)
    cfl = rng.uniform(0.5, 2.5, size=arr["ww"].shape)
    ww = cfl * mass_f / np.abs(m.rdnw[np.minimum(np.arange(nz + 1), nz - 1)][:, None, None] * dts)
    ww *= np.where(rng.random(ww.shape) < 0.5, -1.0, 1.0)
    arr["ww"][...] = ww
    arr["w_work"][...] = rng.uniform(-2.0, 2.0, size=arr["w_work"].shape) * mass_f
    # identity tridiagonal, so the explicit w is the result
    arr["ca"][...] = 0.0
    arr["calpha"][...] = 1.0
    arr["cgamma"][...] = 0.0
    on, off = ({k: v.copy() for k, v in arr.items()} for _ in range(2))
    oracle.piece(m, dataclasses.replace(base, w_damping=1), on, oracle.PIECE_W, rk, dts)
    oracle.piece(m, dataclasses.replace(base, w_damping=0), off, oracle.PIECE_W, rk, dts)
    got = (on["w_work"] - off["w_work"])[1:nz]
    c = np.abs(ww[1:nz] / mass_f[1:nz] * m.rdnw[1:nz, None, None] * dts)
    want = np.where(c > 1.0, -np.sign(arr["w_work"][1:nz]) * 0.3 * (c - 1.0) * mass_f[1:nz] * dts, 0.0)
    assert np.any(c > 1.0) and np.any(c < 1.0)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(want)))
    assert not np.any((on["w_work"] - off["w_work"])[[0, nz]])


def test_log_form_refresh_returns_the_dry_pressure_of_mu_b_plus_mu_pert():
    nx, ny, nz = 5, 4, 16
    eta = 1.0 - (np.arange(nz + 1) / nz) ** 1.3
    m = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=2000.0, dy=2000.0, eta=eta, p_top=PTOP)
    em = 0.5 * (eta[1:] + eta[:-1])
    # dry pressure c3 mu + c4 + p_top with a nonzero c4 (zero at both ends); p' = c3 mu' keeps
    # the w equation's discrete balance rdn (p'_k - p'_k-1) = c1 mu' (c1 = 1 = d c3 / d eta)
    c3h, c3f = em.copy(), eta.copy()
    c4h, c4f = 2.0e3 * em * (1 - em), 2.0e3 * eta * (1 - eta)
    m = dataclasses.replace(m, c3h=c3h, c4h=c4h, c3f=c3f, c4f=c4f)
    Tb, Tt = 260.0, 280.0
    mub = np.full((ny, nx), 9.4e4)
    mut = mub + 600.0  # horizontally uniform, so nothing moves

    def column(mu, T):
        pf = c3f[:, None, None] * mu[None] + c4f[:, None, None] + PTOP
        ph = -R * T * np.log(pf / pf[0])  # isothermal: exact hypsometric geopotential
        phm = c3h[:, None, None] * mu[None] + c4h[:, None, None] + PTOP
        return ph, phm

    phb, pb = column(mub, Tb)
    pht, ph_t = column(mut, Tt)
    theta = Tt * (P0 / ph_t) ** (R / CP)
    st = {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)), "w": np.zeros((nz + 1, ny, nx)),
          "theta": theta, "p_total": ph_t.copy(), "p_perturbation": ph_t - pb, "ph_total": pht,
          "ph_perturbation": pht - phb, "mu_total": mut.copy(), "mu_perturbation": mut - mub,
          "mudf": np.zeros((ny, nx))}
    nl = Namelist(dt_s=6.0, acoustic_substeps=4, hypsometric_opt=2, damp_opt=0)
    arr = oracle.new_arrays(m, st)
    for pid, rk, dts in oracle.step_pieces(nl):
        oracle.piece(m, nl, arr, pid, rk, dts)
        if pid == oracle.PIECE_FINISH:
            break
    # after one stage at rest (mu' unchanged to rounding) the refreshed p is the dry
    # hydrostatic pressure of the column's own mass
    mu_now = arr["mu_total"]
    want = c3h[:, None, None] * mu_now[None] + c4h[:, None, None] + PTOP
    assert np.max(np.abs(arr["mu_total"] - mut)) < 1e-8
    assert np.allclose(arr["p_total"], want, rtol=1e-12, atol=0)
