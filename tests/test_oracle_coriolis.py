"""The oracle's Coriolis terms (large_step_coriolis, dynamics/core/rk_addtend_dry.py:364-483,
added to the coupled u/v tendencies at runtime/operational_mode.py:3052-3058, and the
cosine-Coriolis term of rw_tend, :1818-1821), pinned by closed forms.  No reference
vector covers them (every reference golden case has f = e = 0).

* Linear fields make every face and four-point average exact, so at an interior face
  the u tendency is f m v / msfvx at the face's own position, scaled by msfux / msfuy,
  and the v tendency is -f m u / msfuy scaled by msfvy / msfvx: the signs, the rows and
  columns the averages take, and which map factor divides which momentum.
* The cosine terms average the coupled w of levels k and k + 1: with c1f varying and w
  uniform that is W mu (c1f[k] + c1f[k+1]) / 2, with -e cosa in u and +e sina in v.
* The w term is e (cosa ru - (msftx/msfty) sina rv) with the uniform coupled momenta.
* A whole large step of uniform flow over flat ground, rotating at constant f, is the
  inertial oscillation integrated by WRF's RK3: (u, v) goes to the third-order Taylor
  polynomial of the rotation by f dt, to rounding.  That pins the sign and the stage
  cadence together, and fails a second-order scheme by 1e-6 of U.
"""

import dataclasses

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic


def _flat(nx=12, ny=10, nz=6, dx=2000.0, dy=3000.0, u0=0.0):
    return ic.bubble_3d(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, noise=0.0, amplitude=0.0, u0=u0)


def _rot(m, f=0.0, e=0.0, sina=0.0, cosa=1.0, **more):
    full = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), (m.ny, m.nx)).copy()  # noqa: E731
    return dataclasses.replace(m, f=full(f), e=full(e), sina=full(sina), cosa=full(cosa), **more)


def _with(state, **fields):
    out = {k: np.array(v, copy=True) for k, v in state.items()}
    out.update(fields)
    return out


def _coords(m):
    x_mass = (np.arange(m.nx) + 0.5) * m.dx
    y_mass = (np.arange(m.ny) + 0.5) * m.dy
    x_u = np.arange(m.nx + 1) * m.dx
    y_v = np.arange(m.ny + 1) * m.dy
    return x_mass, y_mass, x_u, y_v


def test_rotation_off_adds_nothing():
    case = _flat(u0=5.0)
    m = case.metrics
    assert not m.rotating
    out = oracle.coriolis_tendency(m, case.state)
    for k in ("u", "v", "w"):
        assert not np.any(out[k])


def test_uniform_flow_gives_f_times_the_crossed_momentum():
    case = _flat()
    m0 = case.metrics
    U, V, f0 = 7.0, -3.0, 1.0e-4
    a, b, c, d = 1.1, 0.9, 1.3, 0.8  # msfux, msfuy, msfvx, msfvy
    m = _rot(m0, f=f0, msfux=np.full_like(m0.msfux, a), msfuy=np.full_like(m0.msfuy, b),
             msfvx=np.full_like(m0.msfvx, c), msfvy=np.full_like(m0.msfvy, d))
    st = _with(case.state, u=np.full_like(case.state["u"], U), v=np.full_like(case.state["v"], V))
    mu = float(case.state["mu_total"][0, 0])
    out = oracle.coriolis_tendency(m, st)
    mass = (m.c1h * mu + m.c2h)[:, None, None]
    want_u = (a / b) * f0 * mass * V / c
    want_v = -(d / c) * f0 * mass * U / b
    assert np.allclose(out["u"], np.broadcast_to(want_u, out["u"].shape), rtol=1e-13, atol=0)
    assert np.allclose(out["v"], np.broadcast_to(want_v, out["v"].shape), rtol=1e-13, atol=0)
    assert not np.any(out["w"])  # e = 0


def test_linear_fields_give_the_face_value_on_a_beta_plane():
    case = _flat(nx=14, ny=11)
    m0 = case.metrics
    x, y, xu, yv = _coords(m0)
    f_of = lambda xx, yy: 1.0e-4 + 2.0e-11 * yy - 5.0e-12 * xx  # noqa: E731
    v_of = lambda xx, yy: 3.0 + 1.0e-4 * xx - 2.0e-4 * yy  # noqa: E731
    u_of = lambda xx, yy: -4.0 + 3.0e-4 * xx + 1.5e-4 * yy  # noqa: E731
    m = _rot(m0, f=f_of(x[None, :], y[:, None]))
    nz = m.nz
    v = np.broadcast_to(v_of(x[None, :], yv[:, None]), (nz, m.ny + 1, m.nx)).copy()
    u = np.broadcast_to(u_of(xu[None, :], y[:, None]), (nz, m.ny, m.nx + 1)).copy()
    st = _with(case.state, u=u, v=v)
    mu = float(case.state["mu_total"][0, 0])
    out = oracle.coriolis_tendency(m, st)
    mass = (m.c1h * mu + m.c2h)[:, None, None]
    # interior u faces: everything at (x_u[i], y[j]); rows j and j+1 of v average to y[j]
    want_u = f_of(xu[None, 1:-1], y[:, None]) * v_of(xu[None, 1:-1], y[:, None])
    got_u = out["u"][:, :, 1:-1] / mass
    assert np.allclose(got_u, np.broadcast_to(want_u, got_u.shape), rtol=1e-11, atol=0)
    # the edge faces take the clamped column x[0] (and x[-1]) for f and for v
    want_u0 = f_of(x[0], y) * v_of(x[0], y)
    assert np.allclose(out["u"][:, :, 0] / mass[:, :, 0], np.broadcast_to(want_u0, (nz, m.ny)), rtol=1e-11)
    # interior v faces at (x[i], y_v[j])
    want_v = -f_of(x[None, :], yv[1:-1, None]) * u_of(x[None, :], yv[1:-1, None])
    got_v = out["v"][:, 1:-1, :] / mass
    assert np.allclose(got_v, np.broadcast_to(want_v, got_v.shape), rtol=1e-11, atol=0)
    want_vn = -f_of(x, y[-1]) * u_of(x, y[-1])
    assert np.allclose(out["v"][:, -1, :] / mass[:, :, 0], np.broadcast_to(want_vn, (nz, m.nx)), rtol=1e-11)


def test_mass_weighting_follows_the_face_mu():
    case = _flat(nx=12, ny=12)
    m0 = case.metrics
    eta_mass = 0.5 * (m0.eta[:-1] + m0.eta[1:])
    m = _rot(m0, f=2.0e-4, c1h=eta_mass, c2h=np.full(m0.nz, 300.0))
    x, y, xu, yv = _coords(m)
    mu_of = lambda yy: 94000.0 + 0.05 * yy  # noqa: E731  mu linear in y
    V = 4.0
    mu = np.broadcast_to(mu_of(y)[:, None], (m.ny, m.nx)).copy()
    st = _with(case.state, mu_total=mu, v=np.full_like(case.state["v"], V), u=np.zeros_like(case.state["u"]))
    out = oracle.coriolis_tendency(m, st)
    want = 2.0e-4 * V * (m.c1h[:, None] * mu_of(y)[None, 1:-1] + m.c2h[:, None])
    assert np.allclose(out["u"][:, 1:-1, 3], want, rtol=1e-12, atol=0)


def test_cosine_terms_average_w_over_levels_k_and_k_plus_1():
    case = _flat(nx=10, ny=9, nz=8)
    m0 = case.metrics
    c1f = np.linspace(1.0, 0.2, m0.nz + 1)
    e0, ca, sa, W = 5.0e-5, 0.8, 0.6, 0.7
    m = _rot(m0, e=e0, cosa=ca, sina=sa, c1f=c1f)
    st = _with(case.state, u=np.zeros_like(case.state["u"]), v=np.zeros_like(case.state["v"]),
               w=np.full_like(case.state["w"], W))
    mu = float(case.state["mu_total"][0, 0])
    out = oracle.coriolis_tendency(m, st)
    rw_mid = W * mu * 0.5 * (c1f[:-1] + c1f[1:])  # (nz,)
    assert np.allclose(out["u"], np.broadcast_to((-e0 * ca * rw_mid)[:, None, None], out["u"].shape), rtol=1e-12)
    assert np.allclose(out["v"], np.broadcast_to((e0 * sa * rw_mid)[:, None, None], out["v"].shape), rtol=1e-12)


def test_w_term_is_e_times_the_rotated_momentum():
    case = _flat(nx=10, ny=9, nz=7)
    m0 = case.metrics
    e0, ca, sa, U, V = 1.2e-4, 0.6, 0.8, 6.0, -2.5
    mx, my = 1.2, 0.9
    m = _rot(m0, e=e0, cosa=ca, sina=sa, msftx=np.full_like(m0.msftx, mx), msfty=np.full_like(m0.msfty, my))
    st = _with(case.state, u=np.full_like(case.state["u"], U), v=np.full_like(case.state["v"], V),
               w=np.zeros_like(case.state["w"]))
    mu = float(case.state["mu_total"][0, 0])
    out = oracle.coriolis_tendency(m, st)
    want = e0 * (ca * mu * U - (mx / my) * sa * mu * V)
    assert np.allclose(out["w"][1:-1], want, rtol=1e-12, atol=0)
    assert not np.any(out["w"][0]) and not np.any(out["w"][-1])


def _taylor_rotation(U, V, th, order):
    # (u, v)' = f (v, -u): the exact map is a rotation by th = f dt; its Taylor
    # polynomial of the given order
    su, sv = U, V
    term_u, term_v = U, V
    for n in range(1, order + 1):
        term_u, term_v = term_v * th / n, -term_u * th / n
        su, sv = su + term_u, sv + term_v
    return su, sv


@pytest.mark.parametrize("ns", [6, 4])
def test_large_step_rotates_uniform_flow_by_the_rk3_polynomial(ns):
    U, V, f0 = 10.0, -4.0, 2.0e-3
    case = _flat(nx=8, ny=6, nz=6, u0=U)
    m = _rot(case.metrics, f=f0)
    nl = dataclasses.replace(case.namelist, acoustic_substeps=ns)
    st = _with(case.state, v=np.full_like(case.state["v"], V))
    out = oracle.run(m, nl, st, 1)
    th = f0 * nl.dt_s
    u3, v3 = _taylor_rotation(U, V, th, 3)
    u2, v2 = _taylor_rotation(U, V, th, 2)
    assert abs(u3 - u2) > 1e-7 * U  # 5e4 times the bar: the test separates RK3 from a 2nd-order scheme
    assert np.max(np.abs(out["u"] - u3)) < 2e-12 * U
    assert np.max(np.abs(out["v"] - v3)) < 2e-12 * U
    # (w is not at rest: uniform flow feels the curvature term, mass (u^2 + v^2) / R)
