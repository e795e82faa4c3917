"""GPU parity of the Coriolis terms (large_step_coriolis, core/rk_addtend_dry.py:364-483,
and the cosine-Coriolis part of rw_tend, operational_mode.py:1818-1821) against the CPU
oracle: terrain with stretched eta, a beta-plane f, e = 2 Omega cos(lat) and a varying
projection rotation (sina, cosa), at the plain FP64 bar (1e-12 of each field's largest
value; observed bitwise)."""

import dataclasses

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu

OMEGA2 = 2.0 * 7.2921e-5


def _rotating(m, lat0=45.0):
    y = (np.arange(m.ny) + 0.5) * m.dy
    x = (np.arange(m.nx) + 0.5) * m.dx
    lat = np.deg2rad(lat0 + (y[:, None] - y.mean()) / 111e3 + 0.0 * x[None, :])
    alpha = np.deg2rad(8.0) * np.sin(2.0 * np.pi * x[None, :] / (m.nx * m.dx)) + 0.0 * y[:, None]
    return dataclasses.replace(m, f=OMEGA2 * np.sin(lat), e=OMEGA2 * np.cos(lat), sina=np.sin(alpha),
                               cosa=np.cos(alpha))


@pytest.mark.parametrize("grid", [(40, 36, 16), (21, 7, 9)])
def test_coriolis_on_terrain_matches_oracle(oracle_lib, grid):
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    nx, ny, nz = grid
    case = ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = _rotating(case.metrics)
    assert m.rotating
    print(f"[coriolis] grid={grid} steps=4", flush=True)
    ref = oracle_lib.run(m, case.namelist, case.state, 4)
    off = oracle_lib.run(case.metrics, case.namelist, case.state, 4)
    assert not np.array_equal(ref["v"], off["v"])
    dyn = Dycore(m, case.namelist)
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


def test_coriolis_mixed_form_and_graph(oracle_lib):
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=24, ny=20, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m = _rotating(case.metrics, lat0=30.0)
    print("[coriolis] mixed form 24x20x12, 6 steps; graph replay 2 steps", flush=True)
    ref = oracle_lib.run(m, case.namelist, case.state, 6)
    dyn = Dycore(m, case.namelist, precision="mixed")
    st = state_to_device(case.state, "cuda")
    dyn.step(st, 6)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    errs = field_errors(got, ref)
    print("mixed field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in FIELDS:
        if k != "mudf":
            assert errs[k] <= 1e-6, (k, errs[k])
    # graph replay of the rotating step equals the eager step bit for bit
    dyn = Dycore(m, case.namelist)
    a = state_to_device(case.state, "cuda")
    b = state_to_device(case.state, "cuda")
    dyn.step(a, 2)
    dyn.capture_graph(b)
    dyn.graph_step(2)
    torch.cuda.synchronize()
    ga, gb = state_to_host(a), state_to_host(b)
    dyn.close()
    for k in FIELDS:
        assert np.array_equal(ga[k], gb[k]), k
