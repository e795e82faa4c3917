"""GPU checks of the mixed-precision form (WD_PREC_MIXED) through the C ABI.

The mixed form is not bit-identical to the oracle: it evaluates the
advection interpolants in FP32 and stores the RK tendencies in FP32.  These
tests hold it to stated error bounds against the FP64 oracle, check that it
really differs from the FP64 form (so the mixed kernels ran), and that its
graph replay equals its eager step.
"""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu

# max|mixed - oracle| / max|oracle| per field after the steps below
MIXED_TOL = 1e-6


def _run_mixed(case, n_steps):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    dyn = Dycore(case.metrics, case.namelist, precision="mixed")
    st = state_to_device(case.state)
    dyn.step(st, n_steps)
    torch.cuda.synchronize()
    out = state_to_host(st)
    dyn.close()
    return out


@pytest.mark.parametrize("name", ["terrain", "bubble"])
def test_mixed_form_error_against_oracle(oracle_lib, name):
    from wrf_gpu_amd import ic

    if name == "terrain":
        case, n = ic.stratified_terrain_3d(nx=40, ny=36, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5), 6
    else:
        case, n = ic.bubble_3d(nx=24, ny=20, nz=12, dx=1000.0, u0=10.0), 10
    print(f"[mixed] case={case.name} steps={n}", flush=True)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, n)
    got = _run_mixed(case, n)
    errs = field_errors(got, ref)
    print("mixed rel err:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in FIELDS:
        assert np.all(np.isfinite(got[k])), k
        assert errs[k] <= MIXED_TOL, (k, errs[k])
    # the mixed kernels ran: the prognostic fields are not the FP64 bits
    assert any(not np.array_equal(got[k], ref[k]) for k in ("u", "v", "w", "theta"))


def test_mixed_graph_replay_matches_mixed_eager():
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=24, ny=20, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5)
    print(f"[mixed-graph] case={case.name}", flush=True)
    eager = Dycore(case.metrics, case.namelist, precision="mixed")
    graphed = Dycore(case.metrics, case.namelist, precision="mixed")
    sa, sb = state_to_device(case.state), state_to_device(case.state)
    graphed.capture_graph(sb)
    for n in (1, 2):
        eager.step(sa, n)
        graphed.graph_step(n)
        torch.cuda.synchronize()
        a, b = state_to_host(sa), state_to_host(sb)
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), k
    eager.close()
    graphed.close()


def test_precision_switch_and_validation():
    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import Dycore

    case = ic.bubble_3d(nx=8, ny=6, nz=5, dx=1000.0)
    d = Dycore(case.metrics, case.namelist, precision="mixed")
    lib = _lib.load()
    assert lib.wd_get_precision(d._handle) == _lib.WD_PREC_MIXED
    assert lib.wd_set_precision(d._handle, 7) == _lib.WD_EINVAL
    assert lib.wd_set_precision(d._handle, _lib.WD_PREC_FP64) == _lib.WD_OK
    assert lib.wd_get_precision(d._handle) == _lib.WD_PREC_FP64
    d.close()
    with pytest.raises(ValueError):
        Dycore(case.metrics, case.namelist, precision="fp16")
