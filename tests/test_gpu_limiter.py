"""GPU parity of the limited theta advection on the final RK stage
(scalar_adv_opt 1 positive-definite, 2 monotonic;
dynamics/flux_advection.py:614-975 applied at
runtime/operational_mode.py:2819-2866) against the CPU oracle: on terrain,
on the 2-D slab (a single row), together with both explicit diffusions, and
through a captured graph, at the plain FP64 path's bar (1e-12 of each
field's largest value), and in the mixed form at its tolerance."""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu


def _with(case, **kw):
    return case.__class__(**{**case.__dict__, "namelist": case.namelist.replace(**kw)})


def _run(case, n_steps, oracle_lib, precision="fp64", graph=False):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    nl = case.namelist
    print(f"[limiter] case={case.name} steps={n_steps} grid={case.metrics.nx}x{case.metrics.ny}x{case.metrics.nz} "
          f"precision={precision} scalar_adv_opt={nl.scalar_adv_opt} graph={graph}", flush=True)
    ref = oracle_lib.run(case.metrics, nl, case.state, n_steps)
    dyn = Dycore(case.metrics, nl, precision=precision)
    st = state_to_device(case.state, "cuda")
    if graph:
        dyn.capture_graph(st)
        dyn.graph_step(n_steps)
    else:
        dyn.step(st, n_steps)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    return got, ref


def _check(got, ref, tol=1e-12):
    errs = field_errors(got, ref)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    print("bitwise equal:", {k: bool(np.array_equal(got[k], ref[k])) for k in FIELDS}, flush=True)
    for k in FIELDS:
        assert errs[k] <= tol, (k, errs[k])


@pytest.mark.parametrize("opt", [1, 2])
@pytest.mark.parametrize("grid", [(40, 36, 16), (21, 7, 9)])
def test_limiter_on_terrain_matches_oracle(oracle_lib, grid, opt):
    from wrf_gpu_amd import ic

    nx, ny, nz = grid
    case = _with(ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5),
                 scalar_adv_opt=opt)
    got, ref = _run(case, 4, oracle_lib)
    off = oracle_lib.run(case.metrics, case.namelist.replace(scalar_adv_opt=0), case.state, 4)
    assert not np.array_equal(ref["theta"], off["theta"])  # the limiter acts
    _check(got, ref)


def test_limiter_on_the_2d_slab_matches_oracle(oracle_lib):
    from wrf_gpu_amd import ic

    case = _with(ic.warm_bubble_2d(), scalar_adv_opt=2)
    got, ref = _run(case, 20, oracle_lib)
    _check(got, ref)


def test_limiter_with_both_diffusions_and_graph_replay(oracle_lib):
    from wrf_gpu_amd import ic

    case = _with(ic.stratified_terrain_3d(nx=24, ny=20, nz=10, dx=3000.0, u_sigma=1.0, sigma=0.5),
                 scalar_adv_opt=1, diff_6th_opt=2, diff_opt=1, km_opt=4)
    got, ref = _run(case, 3, oracle_lib, graph=True)
    _check(got, ref)


def test_limiter_in_the_mixed_form(oracle_lib):
    from wrf_gpu_amd import ic

    case = _with(ic.bubble_3d(nx=24, ny=20, nz=12, dx=1000.0, u0=10.0), scalar_adv_opt=2)
    got, ref = _run(case, 10, oracle_lib, precision="mixed")
    _check(got, ref, tol=1e-6)
