"""GPU runs of the layered forms at their stated workloads, with their error bars.

* Mixed precision on the 64x64x40 warm bubble (dt 6 s, 6 substeps) for 200 large steps,
  against the FP64 oracle: per prognostic field, max|d| / max|ref| and rms(d) / rms(ref)
  both at most 1e-6.  mudf, the carried column mass divergence (the difference of two
  nearly equal column integrals, so its relative rounding error is the largest), is
  reported and held to its own bar.
* The uneven ensemble at its member sizes (129^2, 201^2, 257^2 and 321^2 columns, 50
  levels): the batched FP64 graph equals every member stepped alone bit for bit, and the
  batched mixed form stays within 1e-6 of the batched FP64 form per field.
"""

import numpy as np
import pytest

from tests.parity_util import FIELDS

pytestmark = pytest.mark.gpu

STATE_FIELDS = tuple(k for k in FIELDS if k != "mudf")
MIXED_TOL = 1e-6
MUDF_TOL = 3e-6


def _max_rms(got, ref):
    out = {}
    for k in FIELDS:
        d = got[k] - ref[k]
        mx = float(np.max(np.abs(ref[k])))
        rms = float(np.sqrt(np.mean(ref[k] ** 2)))
        out[k] = (float(np.max(np.abs(d))) / mx if mx else 0.0,
                  float(np.sqrt(np.mean(d ** 2))) / rms if rms else 0.0)
    return out


def test_mixed_form_on_the_64x64x40_bubble_for_200_steps(oracle_lib):
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.bubble_3d()
    steps = 200
    print(f"[configs] mixed form, {case.name}, {steps} steps", flush=True)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, steps)
    dyn = Dycore(case.metrics, case.namelist, precision="mixed")
    st = state_to_device(case.state, "cuda")
    dyn.step(st, steps)
    torch.cuda.synchronize()
    got = state_to_host(st)
    dyn.close()
    errs = _max_rms(got, ref)
    print("max / rms rel err:", {k: f"{a:.2e}/{b:.2e}" for k, (a, b) in errs.items()}, flush=True)
    assert any(not np.array_equal(got[k], ref[k]) for k in FIELDS)  # the form does differ
    for k in STATE_FIELDS:
        assert errs[k][0] <= MIXED_TOL and errs[k][1] <= MIXED_TOL, (k, errs[k])
    assert errs["mudf"][0] <= MUDF_TOL and errs["mudf"][1] <= MUDF_TOL, errs["mudf"]


def test_ensemble_member_sizes_batched_fp64_bitwise_and_mixed_within_bar():
    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device, state_to_host

    cases = [ic.ensemble_member(nx, ny, 50, member=8 * q) for q, (nx, ny) in enumerate(ic.ENSEMBLE_SIZES)]
    steps = 10
    print(f"[configs] ensemble members {[c.name for c in cases]}, {steps} steps", flush=True)
    alone = []
    for c in cases:
        dyn = Dycore(c.metrics, c.namelist)
        st = state_to_device(c.state, "cuda")
        dyn.step(st, steps)
        torch.cuda.synchronize()
        alone.append(state_to_host(st))
        dyn.close()
        del st
    results = {}
    for precision in ("fp64", "mixed"):
        dyns = [Dycore(c.metrics, c.namelist, precision=precision) for c in cases]
        states = [state_to_device(c.state, "cuda") for c in cases]
        ens = Ensemble(list(zip(dyns, states)), branches=4)
        ens.step(steps)
        torch.cuda.synchronize()
        results[precision] = [state_to_host(s) for s in states]
        ens.close()
        for d in dyns:
            d.close()
        del states
        torch.cuda.empty_cache()
    for c, a, b in zip(cases, alone, results["fp64"]):
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), (c.name, k)
    for c, a, b in zip(cases, results["fp64"], results["mixed"]):
        errs = _max_rms(b, a)
        print(c.name, "mixed vs fp64 max / rms:", {k: f"{x:.1e}/{y:.1e}" for k, (x, y) in errs.items()},
              flush=True)
        for k in STATE_FIELDS:
            assert errs[k][0] <= MIXED_TOL and errs[k][1] <= MIXED_TOL, (c.name, k, errs[k])
        assert errs["mudf"][0] <= MUDF_TOL, (c.name, errs["mudf"])
