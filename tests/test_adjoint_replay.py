"""Host replay of the GPU adjoint's reverse launch sequence (wd_step_adjoint in
wrf_gpu_amd/csrc/wrfdyn_adjoint.hip, glue in wrfdyn_adjoint_glue.h).

The executable runs reverse_sequence with every thread of every block of every
reverse kernel, tail threads included, through the same per-thread index maps
and the same tape views as the GPU launches, on a tape filled from the
oracle's traced large step; each tape array is its own allocation and the
build has AddressSanitizer, so any access outside the tape fails the run.
The result must be the gradient oracle's adjoint of the step (orc_vjp)."""

import numpy as np
import pytest

from oracle import oracle
from tests.adjoint_host import replay
from wrf_gpu_amd import ic


@pytest.mark.parametrize("grid", [(12, 10, 8), (21, 7, 9)])
def test_replayed_reverse_sweep_is_the_gradient_oracle(grid):
    case = ic.stratified_terrain_3d(nx=grid[0], ny=grid[1], nz=grid[2], dx=3000.0, u_sigma=1.0, sigma=0.5,
                                    zdamp=10000.0)
    start = oracle.run(case.metrics, case.namelist, case.state, 1)
    rng = np.random.default_rng(5)
    bar = {k: rng.standard_normal(v.shape) for k, v in start.items()}
    got = replay.reverse(case.metrics, case.namelist, start, bar)
    ref, _ = oracle.vjp(case.metrics, case.namelist, start, 1, bar)
    for k in ref:
        scale = max(float(np.abs(ref[k]).max()), 1e-300)
        assert float(np.abs(got[k] - ref[k]).max()) / scale <= 1e-11, k
