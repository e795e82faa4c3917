"""A GPU run resumed from a restart (wrf_gpu_amd.restart) continues bit for
bit: 4 large steps straight through against 2 steps, a restart written from
the device tensors and read back onto the device, then 2 more."""

import numpy as np
import pytest

from wrf_gpu_amd.grid import STATE_FIELDS

pytestmark = pytest.mark.gpu


def test_gpu_run_resumed_from_a_restart_is_bitwise_equal(tmp_path):
    import torch
    from wrf_gpu_amd import ic, restart
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=40, ny=36, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5)
    nl = case.namelist.replace(diff_6th_opt=2, diff_opt=1, km_opt=4, scalar_adv_opt=2)
    print(f"[restart] 40x36x16, 4 steps straight vs 2 + restart + 2", flush=True)
    dyn = Dycore(case.metrics, nl)
    a = state_to_device(case.state)
    dyn.step(a, 4)
    b = state_to_device(case.state)
    dyn.step(b, 2)
    torch.cuda.synchronize()
    p = restart.write_restart(b, case.metrics, nl, 2, tmp_path / "r.npz")
    st, m, nl2, step = restart.read_restart(p, device="cuda")
    dyn2 = Dycore(m, nl2)
    dyn2.step(st, 4 - step)
    torch.cuda.synchronize()
    ra, rb = state_to_host(a), state_to_host(st)
    for k in STATE_FIELDS:
        assert np.array_equal(ra[k], rb[k]), k
    dyn.close()
    dyn2.close()
