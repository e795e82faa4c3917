"""Pin the CPU oracle's operators by mathematics the reference fixes.

Each check is chosen so that a plausible transcription mistake (a wrong
stencil coefficient, sign, index or transposed operand) fails it:

* Thomas sweeps solve the tridiagonal system (scipy.linalg.solve_banded).
* calc_coef_w: interior rows of the implicit w operator sum to one (a
  discrete second-derivative operator annihilates constants), and the top
  rows follow the closed forms of the reference's own test
  (tests/test_m6b0r_jax_top_row_synthetic.py:60-140).
* EOS: the p <-> alpha inversion round-trips and obeys the ideal gas law.
* flux5/flux3 advection converges at 5th order for a smooth periodic field.
* A hydrostatic rest state stays at rest (reference
  tests/test_m6x_vertical_acoustic_oracle.py:159-174).
* A y-uniform 3-D run equals the 2-D slab bit for bit; a transposed initial
  condition gives the transposed result; a mirror-symmetric bubble stays
  mirror symmetric and conserves dry mass.
"""

import dataclasses

import numpy as np
import pytest
import scipy.linalg

from wrf_gpu_amd import ic
from wrf_gpu_amd.grid import Metrics


def test_thomas_matches_solve_banded(oracle_lib):
    rng = np.random.default_rng(0)
    nz = 12
    n = nz + 1
    a = rng.uniform(-0.4, -0.1, n)
    c = rng.uniform(-0.4, -0.1, n)
    b = 1.0 + rng.uniform(0.3, 1.0, n) - a - c
    a[0] = 0.0
    c[nz] = 0.0
    alpha = np.ones(n)
    gamma = np.zeros(n)
    for k in range(1, n):
        alpha[k] = 1.0 / (b[k] - a[k] * gamma[k - 1])
        gamma[k] = c[k] * alpha[k]
    gamma[nz] = 0.0
    rhs = rng.standard_normal(n)
    x = oracle_lib.thomas(a, alpha, gamma, rhs)
    # unknowns x[1..nz] with x[0] = rhs[0] given
    ab = np.zeros((3, nz))
    ab[0, 1:] = c[1:nz]
    ab[1, :] = b[1:]
    ab[2, :-1] = a[2:]
    r = rhs[1:].copy()
    r[0] -= a[1] * rhs[0]
    # the sweep is written as w_k = (rhs_k - a_k w_{k-1}) * alpha_k, i.e. it
    # solves a_k x_{k-1} + b_k x_k + c_k x_{k+1} = rhs_k
    expected = scipy.linalg.solve_banded((1, 1), ab, r)
    assert x[0] == rhs[0]
    np.testing.assert_allclose(x[1:], expected, rtol=1e-12, atol=1e-12)


def _coef_case(nz=10, pure_sigma=True):
    eta = np.linspace(1.0, 0.0, nz + 1)
    m = Metrics.flat(nx=4, ny=3, nz=nz, dx=1000.0, dy=1000.0, eta=eta, p_top=5000.0, pure_sigma=pure_sigma)
    rng = np.random.default_rng(3)
    mut = 95000.0 + rng.uniform(-500, 500, (m.ny, m.nx))
    c2a = 1.0e5 + rng.uniform(-1e4, 1e4, (nz, m.ny, m.nx))
    return m, mut, c2a


def _tridiag_from_coef(a, alpha, gamma):
    b = np.empty_like(alpha)
    b[0] = 1.0 / alpha[0]
    b[1:] = 1.0 / alpha[1:] + a[1:] * gamma[:-1]
    c = gamma / alpha
    return b, c


def test_calc_coef_w_interior_rows_sum_to_one(oracle_lib):
    m, mut, c2a = _coef_case()
    a, alpha, gamma = oracle_lib.calc_coef_w(m, mut, c2a, dts=1.0, epssm=0.1)
    b, c = _tridiag_from_coef(a, alpha, gamma)
    nz = m.nz
    # uniform eta, pure sigma: a_k + b_k + c_k = 1 on interior faces 2..nz-1
    s = a[2:nz] + b[2:nz] + c[2:nz]
    np.testing.assert_allclose(s, 1.0, rtol=0, atol=1e-12)
    # and the operator is not trivial
    assert np.max(np.abs(a[2:nz])) > 1e-3 and np.max(np.abs(c[1:nz])) > 1e-3
    # face 1 has no lower neighbour coupling (a[1] = 0)
    assert np.all(a[1] == 0.0)


def test_calc_coef_w_top_rows_closed_form(oracle_lib):
    """acoustic_wrf.py:808-850: b_top uses c1h(nz-1)*c1f(nz); a_top uses c1f(nz-1)."""

    nz = 8
    eta = np.linspace(1.0, 0.0, nz + 1)
    m = Metrics.flat(nx=4, ny=1, nz=nz, dx=1000.0, dy=1000.0, eta=eta, p_top=5000.0)
    # hybrid-like face coefficients with c1f[nz-1] != c1f[nz] (the reference
    # test's sensitivity guard)
    c1f = np.linspace(1.0, 0.55, nz + 1)
    c2f = (1.0 - c1f) * 2.0e4
    m = dataclasses.replace(m, c1f=c1f, c2f=c2f)
    mut = np.full((1, 4), 90000.0)
    c2a = np.full((nz, 1, 4), 1.0)
    dt, epssm = 6.0, 0.1
    cof = (0.5 * dt * 9.81 * (1.0 + epssm)) ** 2
    a, alpha, gamma = oracle_lib.calc_coef_w(m, mut, c2a, dts=dt, epssm=epssm, top_lid=False)
    b, _ = _tridiag_from_coef(a, alpha, gamma)
    mh = m.c1h[nz - 1] * 90000.0 + m.c2h[nz - 1]
    a_expected = -2.0 * cof * m.rdnw[nz - 1] ** 2 / (mh * (c1f[nz - 1] * 90000.0 + c2f[nz - 1]))
    b_expected = 1.0 + 2.0 * cof * m.rdnw[nz - 1] ** 2 / (mh * (c1f[nz] * 90000.0 + c2f[nz]))
    np.testing.assert_allclose(a[nz], a_expected, rtol=1e-12)
    np.testing.assert_allclose(b[nz], b_expected, rtol=1e-12)
    wrong = -2.0 * cof * m.rdnw[nz - 1] ** 2 / (mh * (c1f[nz] * 90000.0 + c2f[nz]))
    assert abs(wrong - a_expected) > 1e-6 * abs(a_expected)
    # rigid lid zeroes the top lower diagonal
    a_lid, _, _ = oracle_lib.calc_coef_w(m, mut, c2a, dts=dt, epssm=epssm, top_lid=True)
    assert np.all(a_lid[nz] == 0.0)


def test_eos_round_trip_and_ideal_gas(oracle_lib):
    r_d = 287.0
    cp = 7.0 * r_d / 2.0
    for theta in (250.0, 300.0, 420.0):
        for p in (5000.0, 50000.0, 101325.0):
            alt = oracle_lib.alt_from_theta_p(theta, p)
            temp = theta * (p / 1.0e5) ** (r_d / cp)
            assert alt * p == pytest.approx(r_d * temp, rel=1e-13)
            assert oracle_lib.p_from_theta_alt(theta, alt) == pytest.approx(p, rel=1e-13)


def test_flux5_advection_fifth_order(oracle_lib):
    """Uniform transport of sin(2 pi x / L): tendency -> -(ru) dq/dx at 5th order."""

    errs = []
    for nx in (16, 32, 64):
        nz, ny = 4, 1
        lx = 100000.0
        dx = lx / nx
        m = Metrics.flat(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, eta=np.linspace(1, 0, nz + 1), p_top=5000.0)
        mu = np.full((ny, nx), 95000.0)
        u = np.full((nz, ny, nx + 1), 10.0)
        v = np.zeros((nz, ny + 1, nx))
        x = np.arange(nx) * dx  # cell i centred between faces i and i+1
        xc = (np.arange(nx) + 0.5) * dx
        q = np.broadcast_to(np.sin(2 * np.pi * xc / lx), (nz, ny, nx)).copy()
        tend = oracle_lib.advect_scalar(m, u, v, mu, q)
        ru = 95000.0 * 10.0
        exact = -ru * (2 * np.pi / lx) * np.cos(2 * np.pi * xc / lx)
        errs.append(float(np.max(np.abs(tend[0, 0] - exact))) / (ru * 2 * np.pi / lx))
        del x
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(o > 4.6 for o in orders), (errs, orders)


def _rest_case():
    return ic.bubble_3d(nx=8, ny=6, nz=10, dx=1000.0, u0=0.0, noise=0.0, amplitude=0.0,
                        cp=7.0 * 287.0 / 2.0)


def test_hydrostatic_rest_state_stays_at_rest(oracle_lib):
    case = _rest_case()
    s = oracle_lib.run(case.metrics, case.namelist, case.state, 20)
    assert np.max(np.abs(s["w"])) < 1e-10
    assert np.max(np.abs(s["u"])) < 1e-10
    assert np.max(np.abs(s["v"])) < 1e-10
    assert np.max(np.abs(s["ph_perturbation"] - case.state["ph_perturbation"])) < 1e-8
    assert np.max(np.abs(s["theta"] - case.state["theta"])) < 1e-10


def test_y_uniform_3d_equals_2d_slab_bitwise(oracle_lib):
    case = ic.warm_bubble_2d()
    m2 = case.metrics
    ny = 3
    m3 = dataclasses.replace(
        m2, ny=ny, msftx=np.ones((ny, m2.nx)), msfty=np.ones((ny, m2.nx)), msfux=np.ones((ny, m2.nx + 1)),
        msfuy=np.ones((ny, m2.nx + 1)), msfvx=np.ones((ny + 1, m2.nx)), msfvy=np.ones((ny + 1, m2.nx)))
    s3 = {}
    for k, v in case.state.items():
        if k == "v":
            s3[k] = np.zeros((m2.nz, ny + 1, m2.nx))
        else:
            reps = [1] * v.ndim
            reps[-2] = ny
            s3[k] = np.tile(v, reps)
    r2 = oracle_lib.run(m2, case.namelist, case.state, 30)
    r3 = oracle_lib.run(m3, case.namelist, s3, 30)
    for k in r2:
        if k == "v":
            assert np.all(r3[k] == 0.0)
            continue
        for j in range(ny):
            np.testing.assert_array_equal(r3[k][..., j:j + 1, :], r2[k], err_msg=k)


def _transpose_state(s):
    out = {}
    for k, v in s.items():
        if k == "u":
            out["v"] = np.ascontiguousarray(np.swapaxes(v, -1, -2))
        elif k == "v":
            out["u"] = np.ascontiguousarray(np.swapaxes(v, -1, -2))
        else:
            out[k] = np.ascontiguousarray(np.swapaxes(v, -1, -2))
    return out


def test_transposed_ic_gives_transposed_result(oracle_lib):
    case = ic.bubble_3d(nx=14, ny=14, nz=8, dx=1000.0, u0=0.0, noise=0.3, seed=7)
    st = dict(case.state)
    # give the flow a y component too, so u and v both carry signal
    rng = np.random.default_rng(11)
    st["u"] = st["u"] + rng.uniform(-1, 1, st["u"].shape)
    st["v"] = st["v"] + rng.uniform(-1, 1, st["v"].shape)
    ra = oracle_lib.run(case.metrics, case.namelist, st, 5)
    rb = oracle_lib.run(case.metrics, case.namelist, _transpose_state(st), 5)
    rbt = _transpose_state(rb)
    # Transposing swaps the order of some x-then-y sums, so the two runs agree
    # to rounding, not bit for bit (mudf, a divergence of nearly cancelling
    # fluxes, carries ~1e-12); a swapped index or operand shows up at O(1).
    for k in ra:
        scale = float(np.max(np.abs(ra[k]))) or 1.0
        err = float(np.max(np.abs(rbt[k] - ra[k]))) / scale
        assert err < 1e-10, (k, err)


def test_mirror_symmetric_bubble_conserves_mass(oracle_lib):
    case = ic.bubble_3d(nx=16, ny=12, nz=10, dx=1000.0, u0=0.0, noise=0.0)
    s = oracle_lib.run(case.metrics, case.namelist, case.state, 15)
    th = s["theta"]
    assert np.max(np.abs(th - th[:, :, ::-1])) < 1e-11
    assert np.max(np.abs(th - th[:, ::-1, :])) < 1e-11
    u = s["u"]
    assert np.max(np.abs(u + u[:, :, ::-1])) < 1e-11
    m0 = float(np.sum(case.state["mu_total"]))
    m1 = float(np.sum(s["mu_total"]))
    assert abs(m1 - m0) / m0 < 1e-13
    assert np.max(np.abs(s["w"])) > 1e-3  # the bubble does move


def test_deterministic_pow_sin_within_libm_accuracy(oracle_lib):
    """include/wrfdyn_math.h (shared by the kernels and the oracle) stays within
    2 ulp of glibc over the dycore's argument ranges."""

    import math

    rng = np.random.default_rng(5)
    cp = 7.0 * 287.0 / 2.0
    worst = 0.0
    for y in (-(cp - 287.0) / cp, cp / (cp - 287.0)):
        for x in np.concatenate([rng.uniform(0.005, 1.5, 3000), [1.0, 0.5, 0.05, 1.2]]):
            got = oracle_lib.det_pow(float(x), y)
            ref = math.pow(float(x), y)
            worst = max(worst, abs(got - ref) / np.spacing(abs(ref)))
    assert worst <= 2.0, worst
    worst = 0.0
    for x in np.concatenate([rng.uniform(-6.0, 2.0, 3000), [0.0, 1.5707963267948966, 1e-9]]):
        got = oracle_lib.det_sin(float(x))
        ref = math.sin(float(x))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / np.spacing(abs(ref)))
    assert worst <= 2.0, worst


def _surface_case(n, a_u=8.0, b_u=-3.0, a_v=-2.0, b_v=5.0, nz=10, gamma=1.4, length=100000.0, height=500.0):
    """Smooth terrain h = H sin(2 pi x/L) cos(2 pi y/L) on stretched eta, with
    u and v linear in eta at the mass levels (horizontally uniform)."""

    eta = 1.0 - (np.arange(nz + 1) / nz) ** gamma
    d = length / n
    m = Metrics.flat(nx=n, ny=n, nz=nz, dx=d, dy=d, eta=eta, p_top=5000.0)
    x = (np.arange(n) + 0.5) * d
    kx = 2.0 * np.pi / length
    h = height * np.sin(kx * x)[None, :] * np.cos(kx * x)[:, None]
    dhdx = height * kx * np.cos(kx * x)[None, :] * np.cos(kx * x)[:, None]
    dhdy = -height * kx * np.sin(kx * x)[None, :] * np.sin(kx * x)[:, None]
    eta_m = 0.5 * (eta[:-1] + eta[1:])
    u = np.broadcast_to((a_u + b_u * eta_m)[:, None, None], (nz, n, n + 1)).copy()
    v = np.broadcast_to((a_v + b_v * eta_m)[:, None, None], (nz, n + 1, n)).copy()
    us, vs = a_u + b_u * eta[0], a_v + b_v * eta[0]  # the surface is eta = 1
    return m, 9.81 * h, u, v, us * dhdx + vs * dhdy


def test_terrain_surface_w_converges_to_kinematic_condition(oracle_lib):
    """advance_w's terrain surface w (advance_w.py:505-556) is the kinematic
    condition w = u dh/dx + v dh/dy: with u, v extrapolated to the surface by
    cf1..cf3 (exact for profiles linear in eta, stretched eta here) and
    centred terrain differences, it converges to the analytic value at second
    order away from the edge-clamped rows and columns.  A wrong sign, index,
    factor 1/2, swapped direction or extrapolation level breaks the order or
    the limit."""

    errs = []
    for n in (24, 48, 96):
        m, phb0, u, v, exact = _surface_case(n)
        w = oracle_lib.surface_w(m, phb0, u, v)
        inner = (slice(1, n - 1), slice(1, n - 1))
        errs.append(float(np.max(np.abs(w[inner] - exact[inner]))) / float(np.max(np.abs(exact))))
    print("surface w rel errors", errs)
    assert errs[-1] < 2e-3
    for coarse, fine in zip(errs, errs[1:]):
        assert 3.6 < coarse / fine < 4.4


def test_terrain_surface_w_edges_are_one_sided(oracle_lib):
    """The reference clamps the terrain neighbours at the domain edges, so the
    edge columns use a one-sided difference: (h(1) - h(0)) / (2 dx) times the
    surface wind."""

    n = 24
    m, phb0, u, v, _ = _surface_case(n, a_v=0.0, b_v=0.0)
    w = oracle_lib.surface_w(m, phb0, u, v)
    h = phb0 / 9.81
    us = 8.0 - 3.0 * 1.0
    np.testing.assert_allclose(w[:, 0], us * (h[:, 1] - h[:, 0]) / (2.0 * m.dx), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(w[:, -1], us * (h[:, -1] - h[:, -2]) / (2.0 * m.dx), rtol=1e-12, atol=1e-15)
