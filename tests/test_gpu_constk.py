"""GPU parity of the constant-K diffusion (const_nu_m2_s > 0;
dynamics/explicit_diffusion.py:282-331 applied at
runtime/operational_mode.py:2878-2929, with the grid-mean layer thickness of
:2756-2758) against the CPU oracle: on terrain, on the 2-D slab (a single
row: the v faces still diffuse across their two rows, as in the reference),
together with the 6th-order filter (the reference adds the filter first), at
the plain FP64 path's bar (1e-12 of each field's largest value), and in the
mixed form at its tolerance."""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu


def _run(case, n_steps, oracle_lib, precision="fp64"):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    nl = case.namelist
    print(f"[constk] case={case.name} steps={n_steps} grid={case.metrics.nx}x{case.metrics.ny}x{case.metrics.nz} "
          f"precision={precision} diff_6th_opt={nl.diff_6th_opt} nu={nl.const_nu_m2_s}", flush=True)
    ref = oracle_lib.run(case.metrics, nl, case.state, n_steps)
    dyn = Dycore(case.metrics, nl, precision=precision)
    st = state_to_device(case.state, "cuda")
    dyn.step(st, n_steps)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    return got, ref


def _constk(case, nu=75.0, **kw):
    return case.__class__(**{**case.__dict__, "namelist": case.namelist.replace(const_nu_m2_s=nu, **kw)})


@pytest.mark.parametrize("grid", [(40, 36, 16), (21, 7, 9)])
def test_constant_k_on_terrain_matches_oracle(oracle_lib, grid):
    from wrf_gpu_amd import ic

    nx, ny, nz = grid
    case = _constk(ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5), nu=300.0)
    got, ref = _run(case, 6, oracle_lib)
    off = oracle_lib.run(case.metrics, case.namelist.replace(const_nu_m2_s=0.0), case.state, 6)
    assert not np.array_equal(ref["theta"], off["theta"])  # the diffusion acts
    errs = field_errors(got, ref)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    print("bitwise equal:", {k: bool(np.array_equal(got[k], ref[k])) for k in FIELDS}, flush=True)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])


def test_constant_k_on_the_2d_slab_matches_oracle(oracle_lib):
    from wrf_gpu_amd import ic

    case = _constk(ic.warm_bubble_2d())
    got, ref = _run(case, 20, oracle_lib)
    errs = field_errors(got, ref)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])


def test_constant_k_with_the_filter(oracle_lib):
    from wrf_gpu_amd import ic

    case = _constk(ic.bubble_3d(nx=24, ny=20, nz=12, dx=1000.0, u0=10.0), diff_6th_opt=2)
    got, ref = _run(case, 4, oracle_lib)
    errs = field_errors(got, ref)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])


def test_constant_k_in_the_mixed_form(oracle_lib):
    from wrf_gpu_amd import ic

    case = _constk(ic.bubble_3d(nx=24, ny=20, nz=12, dx=1000.0, u0=10.0))
    got, ref = _run(case, 10, oracle_lib, precision="mixed")
    errs = field_errors(got, ref)
    print("mixed field rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in FIELDS:
        assert errs[k] <= 1e-6, (k, errs[k])

