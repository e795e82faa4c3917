"""The oracle's constant-K diffusion (const_nu_m2_s > 0: the flux-form
conservative_constant_k_diffusion_tendency of dynamics/explicit_diffusion.py:282-331
on u, v, w and theta, applied at runtime/operational_mode.py:2878-2929 with the
mean layer thickness of :2756-2758), pinned by closed forms and invariants.
No reference vector covers it.

* With the mass uniform, the horizontal part is K m times the five-point
  Laplacian, whose eigenvalue on a sampled sinusoid is
  -(4 sin^2(a/2) / dx^2 + 4 sin^2(b/2) / dy^2).
* With the mass uniform in the vertical too, a field linear in k has no
  interior tendency and +-m K s / dz^2 in the first and last level (no flux
  through the ends).
* dz is the mean of (ph[k+1] - ph[k]) / g over the grid, at least 1 m.
* In flux form the tendencies sum to zero over each periodic array, u over
  its nx + 1 faces and w over its nz + 1 levels, for any mass.
* The tendency is linear in K: doubling K doubles it bit for bit.
"""

import dataclasses

import numpy as np

from oracle import oracle
from wrf_gpu_amd import ic

G = 9.81


def _flat(nx=16, ny=12, nz=6, dx=3000.0, dy=2000.0):
    return ic.bubble_3d(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, noise=0.0, amplitude=0.0, u0=0.0)


def _with(state, **fields):
    out = {k: np.array(v, copy=True) for k, v in state.items()}
    out.update(fields)
    return out


def _unit_mass(m):
    """Metrics whose level mass is mu itself (c1 = 1, c2 = 0)."""

    return dataclasses.replace(m, c1h=np.ones(m.nz), c2h=np.zeros(m.nz), c1f=np.ones(m.nz + 1),
                               c2f=np.zeros(m.nz + 1))


def test_horizontal_part_is_the_laplacian_eigenvalue():
    case = _flat()
    m, st = case.metrics, case.state
    nx, ny, nz, dx, dy, K = m.nx, m.ny, m.nz, m.dx, m.dy, 75.0
    a, b = 2 * np.pi * 3 / nx, 2 * np.pi * 2 / ny
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    pert = 2.0 * np.sin(a * i + 0.3) * np.sin(b * j + 1.1)
    theta = 300.0 + np.repeat(pert[None], nz, axis=0)
    t, _ = oracle.constk_tendency(m, _with(st, theta=theta), K)
    mu = st["mu_total"]
    assert np.ptp(mu) == 0.0  # flat: the mass is uniform on each level
    lam = -(4 * np.sin(a / 2) ** 2 / dx**2 + 4 * np.sin(b / 2) ** 2 / dy**2)
    mass = m.c1h[:, None, None] * mu[None] + m.c2h[:, None, None]
    want = mass * K * lam * pert[None]
    assert np.allclose(t["theta"], want, rtol=0, atol=1e-11 * np.max(np.abs(want)))


def test_linear_column_has_end_fluxes_only():
    case = _flat()
    m, st = _unit_mass(case.metrics), case.state
    nz, K, s = m.nz, 40.0, 0.5
    theta = np.broadcast_to((300.0 + s * np.arange(nz))[:, None, None], st["theta"].shape).copy()
    t, dz = oracle.constk_tendency(m, _with(st, theta=theta), K)
    mu = st["mu_total"][None]
    end = mu * K * s / dz**2
    assert np.allclose(t["theta"][1:-1], 0.0, rtol=0, atol=1e-12 * np.max(end))
    assert np.allclose(t["theta"][0], end[0], rtol=1e-12, atol=0)
    assert np.allclose(t["theta"][-1], -end[0], rtol=1e-12, atol=0)


def test_dz_is_the_mean_layer_thickness_with_a_floor():
    case = ic.stratified_terrain_3d(nx=20, ny=14, nz=8)
    m, st = case.metrics, case.state
    _, dz = oracle.constk_tendency(m, st, 10.0)
    want = np.mean(np.diff(st["ph_total"], axis=0) / G)
    assert want > 1.0
    assert abs(dz - want) <= 1e-13 * want
    _, dz0 = oracle.constk_tendency(m, _with(st, ph_total=np.zeros_like(st["ph_total"])), 10.0)
    assert dz0 == 1.0


def test_tendencies_sum_to_zero_over_each_periodic_array():
    case = ic.stratified_terrain_3d(nx=20, ny=14, nz=8)
    m, st = case.metrics, case.state
    rng = np.random.default_rng(3)
    fields = {k: st[k] + rng.standard_normal(st[k].shape) for k in ("u", "v", "w", "theta")}
    t, _ = oracle.constk_tendency(m, _with(st, **fields), 75.0)
    for k in ("u", "v", "w", "theta"):
        scale = np.sum(np.abs(t[k]))
        assert scale > 0.0
        assert abs(np.sum(t[k])) <= 1e-12 * scale, k


def test_tendency_is_linear_in_k_bitwise():
    case = ic.stratified_terrain_3d(nx=20, ny=14, nz=8)
    m, st = case.metrics, case.state
    t1, _ = oracle.constk_tendency(m, st, 37.5)
    t2, _ = oracle.constk_tendency(m, st, 75.0)
    for k in t1:
        assert np.array_equal(2.0 * t1[k], t2[k]), k


def test_step_with_constant_k_diffuses_a_bubble():
    case = ic.bubble_3d(nx=16, ny=16, nz=10, dx=1000.0, dy=1000.0, noise=0.0, u0=0.0)
    plain = oracle.run(case.metrics, case.namelist, case.state, 2)
    diffused = oracle.run(case.metrics, case.namelist.replace(const_nu_m2_s=75.0), case.state, 2)
    assert all(np.all(np.isfinite(diffused[k])) for k in diffused)
    pert = lambda s: s["theta"] - s["theta"][:, :1, :1]  # noqa: E731
    # diffusion lowers the bubble's peak excess
    assert np.max(pert(diffused)) < np.max(pert(plain))
