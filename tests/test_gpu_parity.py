"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: per prognostic field, max|gpu - oracle| / max|oracle| <= 1e-12 (FP64,
the tolerance set for the plain FP64 path), on seeded inputs at sizes the
oracle finishes in seconds; the 64x64x40 warm bubble is checked over the full
100 large steps.
"""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors, perturbation_errors

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run_both(case, n_steps, oracle_lib):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    print(f"[parity] case={case.name} steps={n_steps} grid={case.metrics.nx}x{case.metrics.ny}x{case.metrics.nz}",
          flush=True)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, n_steps)
    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state, "cuda")
    dyn.step(st, n_steps)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    return got, ref


def _assert_parity(got, ref, start, tol=TOL):
    errs = field_errors(got, ref)
    perr = perturbation_errors(got, ref, start)
    print("field rel err:", {k: f"{v:.2e}" for k, v in errs.items()})
    print("change rel err:", {k: f"{v:.2e}" for k, v in perr.items()})
    for k in FIELDS:
        assert np.all(np.isfinite(got[k])), k
        assert errs[k] <= tol, (k, errs[k])


def test_warm_bubble_2d(oracle_lib):
    from wrf_gpu_amd import ic

    case = ic.warm_bubble_2d()
    got, ref = _run_both(case, 50, oracle_lib)
    _assert_parity(got, ref, case.state)


def test_bubble_3d_small(oracle_lib):
    from wrf_gpu_amd import ic

    case = ic.bubble_3d(nx=24, ny=20, nz=12, dx=1000.0, u0=10.0)
    got, ref = _run_both(case, 10, oracle_lib)
    _assert_parity(got, ref, case.state)


def test_terrain_small(oracle_lib):
    from wrf_gpu_amd import ic

    case = ic.stratified_terrain_3d(nx=40, ny=36, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5)
    got, ref = _run_both(case, 6, oracle_lib)
    _assert_parity(got, ref, case.state)


@pytest.mark.parametrize("ns", [1, 3])
def test_substep_counts(oracle_lib, ns):
    """ns=1 (stage 2 keeps max(1, ns//2)=1 substep) and an odd ns."""

    from wrf_gpu_amd import ic

    case = ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, u0=5.0, ns=ns, dt=4.0)
    got, ref = _run_both(case, 4, oracle_lib)
    _assert_parity(got, ref, case.state)


def test_minimal_grid(oracle_lib):
    """Smallest supported grid (nx=4, ny=1, nz=4): every stencil wraps."""

    from wrf_gpu_amd import ic

    case = ic.bubble_3d(nx=4, ny=1, nz=4, dx=1000.0, u0=3.0, noise=0.5)
    got, ref = _run_both(case, 3, oracle_lib)
    _assert_parity(got, ref, case.state)


def test_bubble_64x64x40_100_steps(oracle_lib):
    """The 64x64x40 warm-bubble workload, full 100 large steps."""

    from wrf_gpu_amd import ic

    case = ic.bubble_3d()
    got, ref = _run_both(case, 100, oracle_lib)
    _assert_parity(got, ref, case.state)


@pytest.mark.parametrize("column", ["16x16", "8x16", "32x8", "16x8", "march"])
def test_column_kernel_variants(oracle_lib, monkeypatch, column):
    """Every compiled form of the acoustic column kernel (tile shapes of the
    LDS-tiled kernel and the column-march kernel) on a terrain case with the
    damping layer active, at the 50 levels of the 433x433x50 workload (the
    levels-per-thread instantiation it runs); the handle reads
    WRFDYN_COLUMN_TILE at creation."""
    from wrf_gpu_amd import ic

    monkeypatch.setenv("WRFDYN_COLUMN_TILE", column)
    case = ic.stratified_terrain_3d(nx=20, ny=18, nz=50, dx=3000.0, u_sigma=1.0, sigma=0.5)
    got, ref = _run_both(case, 6, oracle_lib)
    _assert_parity(got, ref, case.state)


def test_zero_steps_leave_state_untouched():
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.bubble_3d(nx=12, ny=10, nz=8, dx=1000.0)
    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state)
    dyn.step(st, 0)
    torch.cuda.synchronize()
    got = state_to_host(st)
    for k in case.state:
        assert np.array_equal(got[k], case.state[k]), k
    dyn.close()


def test_streams_and_concurrent_handles(oracle_lib):
    """Steps enqueued on non-default streams, two handles at once on two
    streams, equal the oracle; the handles share nothing."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    cases = [ic.stratified_terrain_3d(nx=24, ny=20, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5),
             ic.bubble_3d(nx=20, ny=16, nz=10, dx=1000.0, u0=8.0)]
    print("[streams] two handles on two streams", flush=True)
    runs = []
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for c, s in zip(cases, streams):
        d = Dycore(c.metrics, c.namelist)
        st = state_to_device(c.state)
        torch.cuda.synchronize()
        d.step(st, 3, stream=s)
        runs.append((d, st))
    torch.cuda.synchronize()
    for (d, st), c in zip(runs, cases):
        ref = oracle_lib.run(c.metrics, c.namelist, c.state, 3)
        got = state_to_host(st)
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k]), k
        d.close()


def test_levels_beyond_lds_tile_limit(oracle_lib):
    """130 levels: more than any LDS tile shape holds (128), so the column
    march must serve it; the autotuner may then only pick the march."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=130, dx=3000.0, u_sigma=1.0, sigma=0.5)
    print("[levels] 12x10x130", flush=True)
    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state)
    cfg = dyn.autotune(st, trial_steps=1)
    assert cfg["column_tile"] == -1
    dyn.step(st, 2)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 2)
    _assert_parity(got, ref, case.state)
