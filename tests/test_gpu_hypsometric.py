"""GPU parity of hypsometric_opt = 2 (the log-pressure specific volume of
WRF calc_p_rho_phi, rk_addtend_dry.py:196-222 and acoustic_wrf.py:394-427)
against the CPU oracle, on terrain with terrain-following dry-pressure
coefficients (c3 = eta), at the plain FP64 path's bar (1e-12 of each
field's largest value)."""

import dataclasses

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [(40, 36, 16), (21, 7, 9)])
def test_log_form_on_terrain_matches_oracle(oracle_lib, grid):
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    nx, ny, nz = grid
    case = ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = case.metrics
    eta_mass = 0.5 * (m.eta[:-1] + m.eta[1:])
    m = dataclasses.replace(m, c3h=eta_mass, c4h=np.zeros(nz), c3f=m.eta.copy(), c4f=np.zeros(nz + 1))
    nl = case.namelist.replace(hypsometric_opt=2)
    print(f"[hypsometric] grid={grid} steps=4", flush=True)
    ref = oracle_lib.run(m, nl, case.state, 4)
    off = oracle_lib.run(m, case.namelist, case.state, 4)
    assert not np.array_equal(ref["p_total"], off["p_total"])
    dyn = Dycore(m, nl)
    st = state_to_device(case.state, "cuda")
    dyn.step(st, 4)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    errs = field_errors(got, ref)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    print("bitwise equal:", {k: bool(np.array_equal(got[k], ref[k])) for k in FIELDS}, flush=True)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])
