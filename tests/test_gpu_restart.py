"""A GPU run resumed from a restart (wrf_gpu_amd.restart) continues bit for
bit: 4 large steps straight through against 2 steps, a restart written from
the device tensors and read back onto the device, then 2 more; both against
the CPU oracle's 4 steps, dry (with both diffusions and the theta limiter) and
moist (with the species' transport)."""

import numpy as np
import pytest

from wrf_gpu_amd.grid import state_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("moist", [False, True])
def test_gpu_run_resumed_from_a_restart_matches_the_oracle(tmp_path, oracle_lib, moist):
    import torch
    from wrf_gpu_amd import ic, restart
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=40, ny=36, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5)
    if moist:
        case = ic.with_moisture(case, moist_adv_opt=2)
        nl = case.namelist.replace(diff_6th_opt=2, scalar_adv_opt=1)
    else:
        nl = case.namelist.replace(diff_6th_opt=2, diff_opt=1, km_opt=4, scalar_adv_opt=2)
    print(f"[restart] 40x36x16 moist={moist}, 4 steps straight vs 2 + restart + 2 vs the oracle", flush=True)
    ref = oracle_lib.run(case.metrics, nl, case.state, 4)
    dyn = Dycore(case.metrics, nl)
    a = state_to_device(case.state)
    dyn.step(a, 4)
    b = state_to_device(case.state)
    dyn.step(b, 2)
    torch.cuda.synchronize()
    p = restart.write_restart(b, case.metrics, nl, 2, tmp_path / "r.npz")
    st, m, nl2, step = restart.read_restart(p, device="cuda")
    assert nl2 == nl
    dyn2 = Dycore(m, nl2)
    dyn2.step(st, 4 - step)
    torch.cuda.synchronize()
    ra, rb = state_to_host(a), state_to_host(st)
    for k in state_fields(nl):
        assert np.array_equal(ra[k], ref[k]), ("straight", k)
        assert np.array_equal(rb[k], ref[k]), ("resumed", k)
    dyn.close()
    dyn2.close()
