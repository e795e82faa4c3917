"""The oracle's 2-D Smagorinsky horizontal diffusion (diff_opt 1, km_opt 4:
the eddy viscosity of dynamics/explicit_diffusion.py:960-1217 and the
coordinate-surface flux divergence of :1243-1365, applied at
runtime/operational_mode.py:2946-2963), pinned by closed forms and
invariants.  No reference vector covers it.

* xkmh = c_s^2 dx dy sqrt(def2), def2 = (D11 - D22)^2 / 4 + D12^2: a pure
  y shear gives the centred difference of the shear, a pure normal strain
  half the difference of the two strains, and a large c_s the cap 10 l.
* With K constant (capped) and the mass uniform, the flux divergence is K m
  times the five-point Laplacian, whose eigenvalue on a sampled sinusoid is
  -(4 sin^2(a/2) / dx^2 + 4 sin^2(b/2) / dy^2).
* In flux form the tendencies of a level sum to zero over the periodic
  domain, for any K and mass; the u and v wrap faces repeat face 0.
* Swapping x and y (u with v) swaps the tendencies.
"""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic

PRANDTL = 1.0 / 3.0


def _flat(nx=16, ny=12, nz=4, dx=3000.0, dy=2000.0):
    return ic.bubble_3d(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, noise=0.0, amplitude=0.0, u0=0.0)


def _with(state, **fields):
    out = {k: np.array(v, copy=True) for k, v in state.items()}
    out.update(fields)
    return out


def _sin2(nlev, ny, nx, a, b, pa, pb, amp=1.0):
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    f = amp * np.sin(a * i + pa) * np.sin(b * j + pb)
    return np.repeat(f[None], nlev, axis=0)


def test_km_of_a_y_shear_is_the_centred_shear():
    case = _flat()
    m, st = case.metrics, case.state
    nx, ny, nz, dx, dy = m.nx, m.ny, m.nz, m.dx, m.dy
    th, U, c_s = 2 * np.pi * 2 / ny, 10.0, 0.25
    u = np.broadcast_to((U * np.sin(th * np.arange(ny) + 0.4))[None, :, None], (nz, ny, nx + 1)).copy()
    _, km = oracle.smag_tendency(m, _with(st, u=u, v=np.zeros((nz, ny + 1, nx))), c_s)
    shear = U * np.cos(th * np.arange(ny) + 0.4) * np.sin(th) / dy
    want = np.broadcast_to((c_s * c_s * dx * dy * np.abs(shear))[None, :, None], km.shape)
    assert np.max(want) < 10.0 * np.sqrt(dx * dy)  # below the cap
    assert np.allclose(km, want, rtol=0, atol=1e-12 * np.max(want))


def test_km_of_a_normal_strain_is_half_the_strain_difference():
    case = _flat()
    m, st = case.metrics, case.state
    nx, ny, nz, dx, dy = m.nx, m.ny, m.nz, m.dx, m.dy
    a, b, U, V, c_s = 2 * np.pi * 3 / nx, 2 * np.pi / ny, 4.0, 7.0, 0.25
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    u = np.broadcast_to((U * np.sin(a * i + 0.3))[None, None, :], (nz, ny, nx + 1)).copy()
    v = np.broadcast_to((V * np.sin(b * j + 1.1))[None, :, None], (nz, ny + 1, nx)).copy()
    u[..., nx] = u[..., 0]
    v[:, ny] = v[:, 0]
    _, km = oracle.smag_tendency(m, _with(st, u=u, v=v), c_s)
    du = 2 * U * np.cos(a * (np.arange(nx) + 0.5) + 0.3) * np.sin(a / 2)
    dv = 2 * V * np.cos(b * (np.arange(ny) + 0.5) + 1.1) * np.sin(b / 2)
    want = c_s * c_s * dx * dy * np.abs(du[None, :] / dx - dv[:, None] / dy)
    assert np.allclose(km, np.broadcast_to(want[None], km.shape), rtol=0, atol=1e-12 * np.max(want))


def test_km_is_capped_at_ten_mixing_lengths():
    case = _flat()
    m, st = case.metrics, case.state
    rng = np.random.default_rng(4)
    s = _with(st, u=rng.standard_normal(st["u"].shape), v=rng.standard_normal(st["v"].shape))
    _, km = oracle.smag_tendency(m, s, 1.0e6)
    assert np.all(km == 10.0 * np.sqrt(m.dx * m.dy))
    _, km_small = oracle.smag_tendency(m, s, 0.25)
    assert np.all(km_small < km) and np.ptp(km_small) > 0


def test_constant_k_is_k_times_mass_times_the_five_point_laplacian():
    """c_s = 1e6 caps K everywhere (every deformation here is far above
    1e-14 /s); the flat case has a uniform mu, so every mass is uniform."""

    case = _flat()
    m, st = case.metrics, case.state
    nx, ny, nz, dx, dy = m.nx, m.ny, m.nz, m.dx, m.dy
    assert np.ptp(st["mu_total"]) == 0.0
    a, b = 2 * np.pi * 3 / nx, 2 * np.pi * 2 / ny
    lap = -(4 * np.sin(a / 2) ** 2 / dx ** 2 + 4 * np.sin(b / 2) ** 2 / dy ** 2)
    u = np.concatenate([_sin2(nz, ny, nx, a, b, 0.3, 0.9, 3.0)] * 2, axis=2)[..., :nx + 1]
    u[..., nx] = u[..., 0]
    v = np.concatenate([_sin2(nz, ny, nx, a, b, 1.3, 0.2, 2.0)] * 2, axis=1)[:, :ny + 1]
    v[:, ny] = v[:, 0]
    w = _sin2(nz + 1, ny, nx, a, b, 0.7, 2.1, 0.5)
    thp = _sin2(nz, ny, nx, a, b, 2.2, 0.4, 1.5)
    tend, km = oracle.smag_tendency(m, _with(st, u=u, v=v, w=w, theta=300.0 + thp), 1.0e6)
    K = 10.0 * np.sqrt(dx * dy)
    assert np.all(km == K)
    mu = st["mu_total"][0, 0]
    mh = (m.c1h * mu + m.c2h)[:, None, None]
    mf = (m.c1f * mu + m.c2f)[:, None, None]
    cases = {"u": (u, K * mh), "v": (v, K * mh), "w": (w, K * mf), "theta": (thp, K / PRANDTL * mh)}
    for name, (f, km_) in cases.items():
        want = km_ * lap * f
        got = tend[name]
        scale = np.max(np.abs(want))
        assert np.allclose(got, want, rtol=0, atol=1e-11 * scale), (name, np.max(np.abs(got - want)) / scale)


@pytest.mark.parametrize("c_s", [0.25, 3.0])
def test_flux_form_sums_to_zero_on_each_level(c_s):
    """Terrain (a varying mu, so varying masses) and random winds and theta;
    c_s 3 caps K over part of the domain."""

    case = ic.stratified_terrain_3d(nx=14, ny=11, nz=5, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m, st = case.metrics, case.state
    nx, ny = m.nx, m.ny
    assert np.ptp(st["mu_total"]) > 0
    rng = np.random.default_rng(8)
    s = _with(st, u=st["u"] + 3 * rng.standard_normal(st["u"].shape),
              v=st["v"] + 3 * rng.standard_normal(st["v"].shape),
              w=rng.standard_normal(st["w"].shape), theta=st["theta"] + rng.standard_normal(st["theta"].shape))
    tend, km = oracle.smag_tendency(m, s, c_s)
    if c_s == 3.0:
        cap = 10.0 * np.sqrt(m.dx * m.dy)
        assert 0 < np.sum(km == cap) < km.size
    core = {"u": tend["u"][..., :nx], "v": tend["v"][:, :ny], "w": tend["w"], "theta": tend["theta"]}
    for name, t in core.items():
        level_sum = np.abs(t.sum(axis=(1, 2)))
        assert np.all(level_sum <= 1e-12 * np.abs(t).sum(axis=(1, 2))), name
        assert np.abs(t).max() > 0
    assert np.array_equal(tend["u"][..., nx], tend["u"][..., 0])
    assert np.array_equal(tend["v"][:, ny], tend["v"][:, 0])


def test_swapping_x_and_y_swaps_the_tendencies():
    case = ic.stratified_terrain_3d(nx=13, ny=13, nz=4, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m, st = case.metrics, case.state
    rng = np.random.default_rng(9)
    s = _with(st, u=st["u"] + rng.standard_normal(st["u"].shape), v=st["v"] + rng.standard_normal(st["v"].shape),
              w=rng.standard_normal(st["w"].shape))
    t = lambda a: np.ascontiguousarray(np.swapaxes(a, -1, -2))
    sw = _with(s, u=t(s["v"]), v=t(s["u"]), w=t(s["w"]), theta=t(s["theta"]), mu_total=t(s["mu_total"]),
               mu_perturbation=t(s["mu_perturbation"]))
    a, ka = oracle.smag_tendency(m, s, 0.25)
    b, kb = oracle.smag_tendency(m, sw, 0.25)
    assert np.allclose(t(kb), ka, rtol=1e-14, atol=0)
    pairs = {"u": "v", "v": "u", "w": "w", "theta": "theta"}
    for name, other in pairs.items():
        scale = np.abs(a[name]).max()
        assert np.allclose(t(b[other]), a[name], rtol=0, atol=1e-13 * scale), name


def test_diffusion_changes_the_step_and_keeps_it_finite():
    case = ic.stratified_terrain_3d(nx=14, ny=11, nz=6, dx=3000.0, u_sigma=1.0, sigma=0.5)
    on = oracle.run(case.metrics, case.namelist.replace(diff_opt=1, km_opt=4), case.state, 3)
    off = oracle.run(case.metrics, case.namelist, case.state, 3)
    for k in on:
        assert np.all(np.isfinite(on[k])), k
    assert not np.array_equal(on["theta"], off["theta"])
    assert not np.array_equal(on["u"], off["u"])
    # the step is still close to the undiffused one
    assert np.max(np.abs(on["theta"] - off["theta"])) < 0.1


@pytest.mark.parametrize("kw", [dict(diff_opt=2, km_opt=4), dict(diff_opt=1, km_opt=0), dict(km_opt=3),
                                dict(c_s=-1.0)])
def test_namelist_outside_the_path_is_rejected(kw):
    case = _flat()
    with pytest.raises(ValueError):
        case.namelist.replace(**kw).validate()
