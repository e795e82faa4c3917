"""The oracle's hypsometric_opt = 2 (WRF calc_p_rho_phi log-pressure
specific volume; rk_addtend_dry.py:196-222, acoustic_wrf.py:394-427), pinned
by the isothermal closed form: with face geopotentials of an isothermal
hydrostatic column, phi = -R T ln(p / p_s) at the face dry pressures
p = c3 mu + c4 + p_top, the log form gives exactly the ideal-gas volume
R T / p at the half-level pressure, so the large-step al is
R T_t / p_h(mu_total) - R T_b / p_h(mub) for a total column at T_t on a base
column at T_b.  The linear form (hypsometric_opt 1) misses it by the
truncation of dp = mu d(eta), which the test shows is far above rounding."""

import dataclasses

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic

R_D = 287.0


def _hybrid(case):
    """The case's metrics with terrain-following dry pressure p = eta mu + p_top
    (c3 = eta, c4 = 0; the flat constructor without the idealized override)."""

    m = case.metrics
    eta_mass = 0.5 * (m.eta[:-1] + m.eta[1:])
    return dataclasses.replace(m, c3h=eta_mass, c4h=np.zeros(m.nz), c3f=m.eta.copy(), c4f=np.zeros(m.nz + 1))


def _isothermal_phi(m, mu, T):
    p = m.c3f[:, None, None] * mu[None] + m.c4f[:, None, None] + m.p_top
    return -R_D * T * np.log(p / p[0:1])


@pytest.mark.parametrize("opt,tol", [(2, 1e-12), (1, None)])
def test_large_step_al_of_isothermal_columns_is_the_closed_form(opt, tol):
    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = _hybrid(case)
    st = {k: np.array(v, copy=True) for k, v in case.state.items()}
    mub = st["mu_total"] - st["mu_perturbation"]
    rng = np.random.default_rng(3)
    mu_pert = 300.0 * rng.standard_normal(mub.shape)
    mu_tot = mub + mu_pert
    T_b, T_t = 270.0, 285.0
    phb = _isothermal_phi(m, mub, T_b)
    pht = _isothermal_phi(m, mu_tot, T_t)
    st["mu_total"], st["mu_perturbation"] = mu_tot, mu_pert
    st["ph_perturbation"] = pht - phb
    st["ph_total"] = phb + (pht - phb)
    arrays = oracle.new_arrays(m, st)
    nl = case.namelist.replace(hypsometric_opt=opt)
    oracle.piece(m, nl, arrays, oracle.PIECE_SETUP, 1, case.namelist.dt_s / 3.0)
    al = arrays["al_ls"]
    ph_h = lambda mu: m.c3h[:, None, None] * mu[None] + m.c4h[:, None, None] + m.p_top
    want = R_D * T_t / ph_h(mu_tot) - R_D * T_b / ph_h(mub)
    err = np.abs(al - want).max() / np.abs(want).max()
    if tol is not None:
        assert err <= tol, err
    else:
        assert err > 1e-6, err  # the linear form is not this closed form


def test_log_form_step_runs_and_differs_from_the_linear_form():
    case = ic.stratified_terrain_3d(nx=14, ny=11, nz=8, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = _hybrid(case)
    a = oracle.run(m, case.namelist.replace(hypsometric_opt=2), case.state, 2)
    b = oracle.run(m, case.namelist, case.state, 2)
    for k in a:
        assert np.all(np.isfinite(a[k])), k
    assert not np.array_equal(a["p_total"], b["p_total"])


def test_log_form_without_the_pressure_coefficients_is_refused():
    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=6, dx=3000.0)
    with pytest.raises(ValueError):
        case.namelist.replace(hypsometric_opt=3).validate()
