"""GPU checks of the adjoint workload's objective and finite differences
(wrf_gpu_amd.sensitivity) against tests/golden/adjoint_fd_257x257x50.json,
which tests/golden/make_adjoint_fd.py wrote from the same plain FP64 path.

The forward path is deterministic, so J and each J(+-eps) reproduce bit for
bit; the central differences are checked for convergence in eps.
"""

import json
import os

import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "adjoint_fd_257x257x50.json")


@pytest.fixture(scope="module")
def golden():
    with open(FIXTURE) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def case():
    from wrf_gpu_amd import ic

    return ic.bubble_3d(nx=257, ny=257, nz=50)


def test_objective_reproduces_bitwise(golden, case):
    from wrf_gpu_amd import sensitivity

    print("[sensitivity] J after 50 steps on 257x257x50", flush=True)
    assert sensitivity.objective(case, golden["n_steps"]) == golden["J"]


def test_central_differences_reproduce_and_converge(golden, case):
    from wrf_gpu_amd import sensitivity

    pts = [p for p in golden["points"] if p["field"] == "theta"]
    big = max(pts, key=lambda p: abs(p["derivative"]))
    small_u = [p for p in golden["points"] if p["field"] == "u"][0]
    print(f"[sensitivity] fd at {big['index']} (theta) and {small_u['index']} (u)", flush=True)
    for p in (big, small_u):
        r = sensitivity.central_differences(case, golden["n_steps"], p["field"], [tuple(p["index"])],
                                            golden["eps"])[0]
        assert (r.j_plus, r.j_minus) == (p["j_plus"], p["j_minus"])
    coarse = sensitivity.central_differences(case, golden["n_steps"], "theta", [tuple(big["index"])],
                                             10 * golden["eps"])[0]
    rel = abs(coarse.derivative - big["derivative"]) / abs(big["derivative"])
    print(f"[sensitivity] dJ/dtheta eps 1e-5 vs 1e-6: rel diff {rel:.2e}", flush=True)
    assert rel < 1e-3
