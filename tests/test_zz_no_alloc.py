"""GPU check of wd_step_noalloc, the entry point that allocates nothing (its
metrics and workspace live in the library's fixed device scratch), through the
C ABI: two calls in a row, the second on a smaller grid, each bit for bit the
handle path's wd_step on the same inputs; a grid too large for the scratch is
refused before any device work."""

import dataclasses

import numpy as np
import pytest

from tests.parity_util import FIELDS

pytestmark = pytest.mark.gpu


def _handle_path(case, steps):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state, "cuda")
    dyn.step(st, steps)
    torch.cuda.synchronize()
    out = state_to_host(st)
    dyn.close()
    return out


def test_noalloc_step_equals_the_handle_step_on_two_calls():
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import state_to_device, state_to_host, step_noalloc

    big = ic.bubble_3d()  # 64x64x40, the largest workload the scratch holds
    small = ic.stratified_terrain_3d(nx=24, ny=20, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5)
    small = dataclasses.replace(small, namelist=small.namelist.replace(diff_6th_opt=2))
    for case, steps in ((big, 2), (small, 3)):
        print(f"[noalloc] {case.name}, {steps} steps", flush=True)
        want = _handle_path(case, steps)
        st = state_to_device(case.state, "cuda")
        step_noalloc(st, case.namelist, case.metrics, steps)
        torch.cuda.synchronize()
        got = state_to_host(st)
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (case.name, k)


def test_noalloc_refuses_a_grid_beyond_the_scratch():
    import torch
    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import state_to_device, step_noalloc

    case = ic.bubble_3d(nx=96, ny=96, nz=40, noise=0.0)
    st = state_to_device(case.state, "cuda")
    before = {k: v.clone() for k, v in st.items()}
    with pytest.raises(_lib.WrfdynError, match="scratch"):
        step_noalloc(st, case.namelist, case.metrics, 1)
    torch.cuda.synchronize()
    for k in st:
        assert torch.equal(st[k], before[k]), k
