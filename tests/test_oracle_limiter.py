"""The oracle's positive-definite (scalar_adv_opt 1) and monotonic (2)
limited scalar advection (dynamics/flux_advection.py:614-975), which the
reference applies to theta - 300 on the final RK stage
(runtime/operational_mode.py:2819-2866), pinned by the properties that
define it.  No reference vector covers it.

* Where nothing needs limiting (a smooth positive field inside its local
  extrema), the limited tendency is the plain h5/v3 one.
* Positive definite: with a field that has zeros and spikes, the explicit
  update by dt of the coupled value stays non-negative, while the plain
  scheme's goes negative.
* Monotonic: in a flow that keeps the mass, the update stays within each
  cell's local extrema of the origin field (itself and its six neighbours),
  which the plain scheme overshoots.
* Conservative: the tendency weighted by the layer thickness sums to zero.
"""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic


def _case(nx=16, ny=12, nz=8):
    return ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=2.0, sigma=0.5)


def _spiky(shape, rng):
    """Non-negative, mostly zero, with isolated spikes and a smooth part."""
    q = np.zeros(shape)
    mask = rng.random(shape) < 0.08
    q[mask] = rng.uniform(0.5, 2.0, size=int(mask.sum()))
    return q


def _update(m, q, mu_old, tend, dt):
    mass_old = m.c1h[:, None, None] * mu_old[None] + m.c2h[:, None, None]
    return mass_old * q + dt * tend, mass_old


def _wind(st, rng, scale):
    u = st["u"] + scale * rng.standard_normal(st["u"].shape)
    v = st["v"] + scale * rng.standard_normal(st["v"].shape)
    u[..., -1] = u[..., 0]
    v[:, -1] = v[:, 0]
    return u, v


@pytest.mark.parametrize("opt", [1, 2])
def test_smooth_field_inside_its_extrema_is_not_limited(opt):
    case = _case()
    m, st = case.metrics, case.state
    nz, ny, nx = m.nz, m.ny, m.nx
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    q = 5.0 + np.sin(2 * np.pi * i / nx) * np.cos(2 * np.pi * j / ny) + 0.1 * k
    mu = st["mu_total"]
    plain = oracle.advect_scalar(m, st["u"], st["v"], mu, q)
    # a tiny dt: the fluxes move almost nothing, so no cell leaves its bounds
    # and the limiter keeps every scale at 1 (monotonic: wherever the cell is
    # not itself a local extremum, which a sinusoid's crest rows are, so
    # compare where the bound is not tight)
    lim = oracle.advect_limited(m, opt, st["u"], st["v"], mu, q, q, mu, 1.0e-3)
    scale = np.abs(plain).max()
    if opt == 1:
        assert np.allclose(lim, plain, rtol=0, atol=1e-11 * scale)
    else:
        close = np.isclose(lim, plain, rtol=0, atol=1e-11 * scale)
        assert close.mean() > 0.5


def test_positive_definite_update_stays_non_negative():
    case = _case()
    m, st = case.metrics, case.state
    rng = np.random.default_rng(3)
    u, v = _wind(st, rng, 8.0)
    mu = st["mu_total"]
    q = _spiky((m.nz, m.ny, m.nx), rng)
    dt = 60.0
    lim = oracle.advect_limited(m, 1, u, v, mu, q, q, mu, dt)
    plain = oracle.advect_scalar(m, u, v, mu, q)
    new_lim, mass = _update(m, q, mu, lim, dt)
    new_plain, _ = _update(m, q, mu, plain, dt)
    scale = np.abs(mass * q).max()
    assert new_plain.min() < -1e-3 * scale  # the plain scheme undershoots
    assert new_lim.min() >= -1e-12 * scale
    assert not np.allclose(lim, plain)


def test_monotonic_update_stays_within_local_extrema():
    """The bound holds where the transport does not change the mass (the
    low-order update is then itself monotone): a uniform wind over a flat,
    uniform-mass column."""

    case = ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, noise=0.0, amplitude=0.0, u0=0.0)
    m, st = case.metrics, case.state
    rng = np.random.default_rng(4)
    u = np.full(st["u"].shape, 6.0)
    v = np.full(st["v"].shape, -4.0)
    mu = st["mu_total"]
    assert np.ptp(mu) == 0.0
    q = 1.0 + _spiky((m.nz, m.ny, m.nx), rng)
    dt = 60.0
    lim = oracle.advect_limited(m, 2, u, v, mu, q, q, mu, dt)
    plain = oracle.advect_scalar(m, u, v, mu, q)
    qmax, qmin = q.copy(), q.copy()
    for ax in (2, 1):
        for sh in (1, -1):
            r = np.roll(q, sh, axis=ax)
            qmax, qmin = np.maximum(qmax, r), np.minimum(qmin, r)
    up = np.concatenate([q[:1], q[:-1]])
    dn = np.concatenate([q[1:], q[-1:]])
    qmax, qmin = np.maximum(qmax, np.maximum(up, dn)), np.minimum(qmin, np.minimum(up, dn))
    new_lim, mass = _update(m, q, mu, lim, dt)
    new_plain, _ = _update(m, q, mu, plain, dt)
    tol = 1e-12 * np.abs(mass * q).max()
    over = lambda new: max((new - qmax * mass).max(), (qmin * mass - new).max())
    assert over(new_plain) > 1e-3 * np.abs(mass).max()  # the plain scheme leaves the bounds
    assert over(new_lim) <= tol


@pytest.mark.parametrize("opt", [1, 2])
def test_limited_tendency_is_conservative(opt):
    case = _case()
    m, st = case.metrics, case.state
    rng = np.random.default_rng(5)
    u, v = _wind(st, rng, 8.0)
    mu = st["mu_total"]
    q = _spiky((m.nz, m.ny, m.nx), rng) + 0.3
    lim = oracle.advect_limited(m, opt, u, v, mu, q, q, mu, 60.0)
    dnw = -1.0 / m.rdnw  # layer thickness in eta, positive
    total = (lim * np.abs(dnw)[:, None, None]).sum()
    assert abs(total) <= 1e-12 * (np.abs(lim) * np.abs(dnw)[:, None, None]).sum()


@pytest.mark.parametrize("opt", [1, 2])
def test_limiter_on_the_final_stage_changes_the_step(opt):
    case = _case(nx=14, ny=11, nz=6)
    on = oracle.run(case.metrics, case.namelist.replace(scalar_adv_opt=opt), case.state, 2)
    off = oracle.run(case.metrics, case.namelist, case.state, 2)
    for k in on:
        assert np.all(np.isfinite(on[k])), k
    assert not np.array_equal(on["theta"], off["theta"])
    assert np.max(np.abs(on["theta"] - off["theta"])) < 0.5


def test_piece_api_refuses_the_limiter():
    case = _case(nx=12, ny=10, nz=8)
    nl = case.namelist.replace(scalar_adv_opt=1)
    arrays = oracle.new_arrays(case.metrics, case.state)
    with pytest.raises(RuntimeError):
        oracle.piece(case.metrics, nl, arrays, oracle.PIECE_UV, 1, 1.0)
