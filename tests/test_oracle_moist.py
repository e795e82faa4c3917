"""The oracle's moist large step: the six species qv, qc, qr, qi, qs, qg couple into the
dynamics through the WRF moisture factors (dynamics/acoustic_wrf.py:303-319,
core/advance_w.py:195-290, core/rk_addtend_dry.py:278, :359-360) and, with moist_adv_opt,
are transported by the flux-form scalar advection with the large-step update
(runtime/operational_mode.py:3162-3316, :3379-3393, :3489-3495).  No reference vector
holds a moist step, so the pins are:

* the factors in closed form for uniform and linear total moisture, with every species
  counted, the edge faces clamped and the raw w-face factor 0 at the surface and lid
  (the reference's own factor tests, tests/test_v014_moist_cqw_pressure_dynamics.py);
* a state whose species are all zero steps bit for bit as the dry state, for each
  moist_adv_opt (the reference's inertness test, there on pg_buoy_w alone);
* the water loading: for a column at rest with uniform qv = Q, the first large step's w
  differs from the dry step's by -cq2 (w_dry + g dt mu_b / mu) + O(dt^2), cq2 = Q / (1 + Q):
  the dry pressure imbalance scaled by cq1 and the loading -g cq2 of rw_tend's moist form;
* the horizontal pressure gradient: u of a resting bubble after a short step is the dry
  u times cqu = 1 / (1 + Q) to O(dt);
* the implicit w operator: each interior row's departure from the identity (a, b - 1, c)
  is the dry row's times the cq1 of that row's own face;
* transport: moist_adv_opt 0 leaves every species bit for bit as it was; with 1 or 2 the
  column-mass-weighted total of each species, sum |dnw| (c1h mu + c2h) q, is conserved to
  rounding over doubly periodic flow, and the positive-definite option keeps a sharp
  non-negative blob non-negative.

The moisture factor of the explicit part of the implicit w equation and of the acoustic
u/v pressure gradient enter a step only at second order in dt: no closed form above
isolates them, so those two uses stay parity unpinned.
"""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic

G = 9.81
SPECIES = oracle.MOIST_SPECIES


def _dry_moist(state, **q):
    out = {k: np.array(v, copy=True) for k, v in state.items()}
    for name in SPECIES:
        out[name] = np.array(q[name], copy=True) if name in q else np.zeros_like(out["theta"])
    return out


def _uniform_species():
    # binary fractions so every partial sum is exact and each species is visible
    return {"qv": 2.0 ** -7, "qc": 2.0 ** -10, "qr": 2.0 ** -11, "qi": 2.0 ** -12, "qs": 2.0 ** -13,
            "qg": 2.0 ** -14}


def test_factors_uniform_moisture():
    case = ic.bubble_3d(nx=8, ny=6, nz=6, dx=2000.0, noise=0.0, amplitude=0.0, u0=0.0)
    m = case.metrics
    vals = _uniform_species()
    st = _dry_moist(case.state, **{k: np.full_like(case.state["theta"], v) for k, v in vals.items()})
    out = oracle.moist_factors(m, st)
    Q = sum(vals.values())
    assert np.all(out["qtot"] == Q)
    assert np.all(out["cqu"] == 1.0 / (1.0 + Q))
    assert np.all(out["cqv"] == 1.0 / (1.0 + Q))
    assert np.all(out["cqw"][1:m.nz] == Q)
    assert not np.any(out["cqw"][0]) and not np.any(out["cqw"][m.nz])


def test_factors_linear_moisture_and_edge_clamping():
    case = ic.bubble_3d(nx=9, ny=7, nz=6, dx=2000.0, noise=0.0, amplitude=0.0, u0=0.0)
    m = case.metrics
    nz, ny, nx = m.nz, m.ny, m.nx
    k = np.arange(nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    a, bx, by, bz = 0.004, 0.0005, 0.00025, 0.001
    qtot = a + bx * i + by * j + bz * k + 0.0 * (i + j + k)
    # split over the species in fixed fractions; the factors see only the sum
    parts = {"qv": 0.5, "qc": 0.25, "qr": 0.125, "qi": 0.0625, "qs": 0.03125, "qg": 0.03125}
    st = _dry_moist(case.state, **{s: f * qtot for s, f in parts.items()})
    out = oracle.moist_factors(m, st)
    assert np.allclose(out["qtot"], qtot, rtol=1e-15, atol=0)
    qt = out["qtot"]
    # u faces: the mean of the two mass points beside the face, x-linear inside
    iu = np.arange(1, nx)[None, None, :]
    want_u_in = 1.0 / (1.0 + (a + bx * (iu - 0.5) + by * j + bz * k))
    assert np.allclose(out["cqu"][:, :, 1:nx], want_u_in, rtol=1e-14, atol=0)
    assert np.array_equal(out["cqu"][:, :, 0], 1.0 / (1.0 + 0.5 * (qt[:, :, 0] + qt[:, :, 0])))
    assert np.array_equal(out["cqu"][:, :, nx], 1.0 / (1.0 + 0.5 * (qt[:, :, nx - 1] + qt[:, :, nx - 1])))
    jv = np.arange(1, ny)[None, :, None]
    want_v_in = 1.0 / (1.0 + (a + bx * i + by * (jv - 0.5) + bz * k))
    assert np.allclose(out["cqv"][:, 1:ny, :], want_v_in, rtol=1e-14, atol=0)
    assert np.array_equal(out["cqv"][:, 0, :], 1.0 / (1.0 + qt[:, 0, :]))
    assert np.array_equal(out["cqv"][:, ny, :], 1.0 / (1.0 + qt[:, ny - 1, :]))
    # w faces: the raw face mean, z-linear inside, 0 at the surface and the lid
    kf = np.arange(1, nz)[:, None, None]
    assert np.allclose(out["cqw"][1:nz], a + bx * i + by * j + bz * (kf - 0.5), rtol=1e-14, atol=0)
    assert not np.any(out["cqw"][0]) and not np.any(out["cqw"][nz])


@pytest.mark.parametrize("opt", [0, 1, 2])
def test_zero_moisture_steps_as_the_dry_state_bitwise(opt):
    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=3000.0)
    m, nl = case.metrics, case.namelist
    dry = oracle.run(m, nl, case.state, 2, threads=4)
    moist = oracle.run(m, nl.replace(moist=True, moist_adv_opt=opt), _dry_moist(case.state), 2, threads=4)
    for key, v in dry.items():
        assert np.array_equal(moist[key], v), key
    for q in SPECIES:
        assert not np.any(moist[q]), q


def test_water_loading_is_the_first_order_w_tendency():
    # a column at rest whose theta ramp leaves the pressure out of balance: the dry step's
    # w is dt times that imbalance, g rdn dp' / mu at first order (the stage's own
    # diagnosed p'); moisture scales the imbalance by cq1 and adds the loading -g cq2, so
    # w_moist - w_dry = -cq2 (w_dry + g dt (c1f mu_b + c2f) / (c1f mu + c2f)) + O(dt^2)
    Q = 0.01
    cq2 = Q / (1.0 + Q)
    errs = []
    for dt in (0.02, 0.002):
        case = ic.bubble_3d(nx=6, ny=4, nz=8, dx=2000.0, noise=0.0, amplitude=0.0, u0=0.0, dt=dt, ns=2)
        m, nl = case.metrics, case.namelist
        nz = m.nz
        state = dict(case.state)
        state["theta"] = state["theta"] + np.linspace(0.0, 3.0, nz)[:, None, None]
        dry = oracle.run(m, nl, state, 1, threads=4)
        moist = oracle.run(m, nl.replace(moist=True), _dry_moist(state, qv=np.full_like(state["theta"], Q)), 1,
                           threads=4)
        mu = state["mu_total"]
        mub = mu - state["mu_perturbation"]
        c1f, c2f = m.c1f[1:nz, None, None], m.c2f[1:nz, None, None]
        assert np.max(np.abs(dry["w"][1:nz])) > 0.05 * dt  # the imbalance is of the loading's order
        want = -cq2 * (dry["w"][1:nz] + G * dt * (c1f * mub + c2f) / (c1f * mu + c2f))
        got = moist["w"][1:nz] - dry["w"][1:nz]
        errs.append(np.max(np.abs(got / want - 1.0)))
        # horizontally uniform loading moves no air sideways
        assert not np.any(moist["u"]) and not np.any(moist["v"])
    # the remainder is second order: a hundredfold smaller for a tenth of dt
    assert errs[1] < 1e-6 and errs[1] < 0.02 * errs[0], errs


def test_horizontal_pressure_gradient_scaled_by_cqu():
    Q = 0.01
    for dt, tol in ((0.01, 2e-6), (0.001, 2e-7)):
        case = ic.bubble_3d(nx=12, ny=10, nz=8, dx=2000.0, noise=0.0, amplitude=2.0, u0=0.0, dt=dt, ns=2)
        m, nl = case.metrics, case.namelist
        dry = oracle.run(m, nl, case.state, 1, threads=4)
        moist = oracle.run(m, nl.replace(moist=True), _dry_moist(case.state, qv=np.full_like(case.state["theta"], Q)),
                           1, threads=4)
        for key in ("u", "v"):
            sel = np.abs(dry[key]) > 1e-3 * np.abs(dry[key]).max()
            ratio = moist[key][sel] / dry[key][sel] * (1.0 + Q)
            assert np.max(np.abs(ratio - 1.0)) < tol, (key, dt)


def test_implicit_w_rows_scale_by_the_face_cq1():
    case = ic.stratified_terrain_3d(nx=6, ny=5, nz=10, dx=3000.0)
    m = case.metrics
    rng = np.random.default_rng(3)
    nz = m.nz
    mut = case.state["mu_total"]
    c2a = 1.0e5 * (1.0 + 0.2 * rng.uniform(-1, 1, (nz, m.ny, m.nx)))
    qtot = 0.02 * rng.uniform(0, 1, (nz, m.ny, m.nx))
    raw = np.zeros((nz + 1, m.ny, m.nx))
    raw[1:nz] = 0.5 * (qtot[1:] + qtot[:-1])
    cq1 = 1.0 / (1.0 + raw)

    def rows(cqw):
        a, alpha, gamma = oracle.calc_coef_w(m, mut, c2a, dts=2.0, epssm=0.1, cqw_raw=cqw)
        b = np.zeros_like(a)
        c = np.zeros_like(a)
        for k in range(1, nz):
            b[k] = 1.0 / alpha[k] + a[k] * gamma[k - 1]
            c[k] = gamma[k] / alpha[k]
        return a[1:nz], b[1:nz], c[1:nz]

    ad, bd, cd = rows(None)
    am, bm, cm = rows(raw)
    f = cq1[1:nz]
    assert np.allclose(am, f * ad, rtol=1e-12, atol=0)
    assert np.allclose(cm, f * cd, rtol=1e-12, atol=0)
    assert np.allclose(bm - 1.0, f * (bd - 1.0), rtol=1e-10, atol=0)
    # zero moisture gives the dry rows bit for bit
    a0, b0, c0 = rows(np.zeros_like(raw))
    assert np.array_equal(a0, ad) and np.array_equal(c0, cd)


def test_transport_off_leaves_species_untouched():
    case = ic.with_moisture(ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=3000.0), moist_adv_opt=0)
    out = oracle.run(case.metrics, case.namelist, case.state, 2, threads=4)
    for q in SPECIES:
        assert np.array_equal(out[q], case.state[q]), q
    # the coupling still acts on the dynamics
    dry = oracle.run(case.metrics, case.namelist.replace(moist=False), {k: v for k, v in case.state.items()
                                                                         if k not in SPECIES}, 2, threads=4)
    assert not np.array_equal(out["w"], dry["w"])


def _column_mass_total(m, state, q):
    mass = m.c1h[:, None, None] * state["mu_total"][None] + m.c2h[:, None, None]
    return float(np.sum(np.abs(m.dnw)[:, None, None] * mass * state[q]))


@pytest.mark.parametrize("opt", [1, 2])
def test_species_mass_conserved(opt):
    case = ic.with_moisture(ic.stratified_terrain_3d(nx=14, ny=12, nz=8, dx=3000.0, u0=15.0), moist_adv_opt=opt)
    m = case.metrics
    st = case.state
    before = {q: _column_mass_total(m, st, q) for q in SPECIES}
    for _ in range(2):
        st = oracle.run(m, case.namelist, st, 1, threads=4)
        for q in SPECIES:
            after = _column_mass_total(m, st, q)
            assert abs(after - before[q]) <= 1e-13 * before[q] * 50, (q, after, before[q])
    # the species did move
    assert not np.array_equal(st["qc"], case.state["qc"])


def test_positive_definite_keeps_a_sharp_blob_non_negative():
    base = ic.bubble_3d(nx=16, ny=6, nz=8, dx=1000.0, noise=0.0, amplitude=2.0, u0=20.0, dt=4.0, ns=4)
    nz, ny, nx = base.metrics.nz, base.metrics.ny, base.metrics.nx
    blob = np.zeros((nz, ny, nx))
    blob[2:4, 2:4, 5:7] = 1.0e-3  # a box: the plain h5/v3 fluxes undershoot at its edges
    st = _dry_moist(base.state, qc=blob)
    pd = oracle.run(base.metrics, base.namelist.replace(moist=True, moist_adv_opt=1), st, 3, threads=4)
    assert pd["qc"].min() >= -1e-18
    assert pd["qc"].max() > 0.0
