"""GPU checks of the adjoint (wd_step_adjoint and the torch.autograd path of
wrf_gpu_amd.adjoint) against the gradient oracle (oracle.vjp) on small grids,
and against the committed central differences at 257x257x50.

The adjoint sums contributions with atomic adds, so it agrees with the
oracle's reverse sweep to rounding: 1e-10 of each field's largest entry.

wd_step_adjoint is switched off: its first GPU run (the first test below)
ended in a memory access fault whose cause is not yet found, so every test
that would launch its kernels is skipped until the cause is shown.  The
adjoint bodies themselves are checked on the host (test_adjoint_pieces.py).
"""

import json
import os

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
SWITCHED_OFF = ("wd_step_adjoint is switched off: its first GPU run ended in a memory access fault "
                "whose cause is not yet found")
FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "adjoint_fd_257x257x50.json")


def _small(nx=12, ny=10, nz=8):
    from wrf_gpu_amd import ic

    case = ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5, zdamp=10000.0)
    return case, oracle.run(case.metrics, case.namelist, case.state, 2)


def _check(got: dict, ref: dict, what: str):
    for k in ref:
        scale = max(float(np.abs(ref[k]).max()), 1e-300)
        err = float(np.abs(got[k] - ref[k]).max()) / scale
        print(f"[adjoint] {what} {k:16s} rel {err:.2e}", flush=True)
        assert err <= TOL, (what, k, err)


@pytest.mark.skip(reason=SWITCHED_OFF)
@pytest.mark.parametrize("grid", [(12, 10, 8), (21, 7, 9)])
def test_step_adjoint_matches_gradient_oracle(grid):
    """Two large steps: the state at the start of each is kept, and
    wd_step_adjoint runs from the last step back to the first."""

    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case, start = _small(*grid)
    print(f"[adjoint] grid {grid}, 2 steps", flush=True)
    rng = np.random.default_rng(5)
    bar = {k: rng.standard_normal(v.shape) for k, v in start.items()}
    ref, final = oracle.vjp(case.metrics, case.namelist, start, 2, bar)
    dyn = Dycore(case.metrics, case.namelist)
    s0 = state_to_device(start)
    s1 = {k: v.clone() for k, v in s0.items()}
    dyn.step(s1, 1)
    b = state_to_device(bar)
    dyn.step_adjoint(s1, b)
    dyn.step_adjoint(s0, b)
    torch.cuda.synchronize()
    _check(state_to_host(b), ref, f"grid {grid}")
    # the starting states are not modified
    assert all(np.array_equal(state_to_host(s0)[k], start[k]) for k in start)
    assert dyn.adjoint_tape_bytes > 0
    dyn.close()


@pytest.mark.skip(reason=SWITCHED_OFF)
def test_autograd_top_half_w2_matches_gradient_oracle():
    """J = sum(w**2) over the top half of the w levels after three steps,
    differentiated by torch.autograd through wrf_gpu_amd.adjoint.advance."""

    import torch
    from wrf_gpu_amd import adjoint
    from wrf_gpu_amd.dycore import Dycore, state_to_device

    case, start = _small()
    nz, n = case.metrics.nz, 3
    final = oracle.run(case.metrics, case.namelist, start, n)
    bar_w = np.zeros_like(final["w"])
    bar_w[(nz + 1) // 2:] = 2.0 * final["w"][(nz + 1) // 2:]
    ref, _ = oracle.vjp(case.metrics, case.namelist, start, n, {"w": bar_w})
    dyn = Dycore(case.metrics, case.namelist)
    x = {k: v.requires_grad_(True) for k, v in state_to_device(start).items()}
    out = adjoint.advance(dyn, x, n)
    assert np.array_equal(out["w"].detach().cpu().numpy(), final["w"])
    J = (out["w"][(nz + 1) // 2:] ** 2).sum()
    J.backward()
    torch.cuda.synchronize()
    _check({k: x[k].grad.cpu().numpy() for k in x}, ref, "autograd J")
    dyn.close()


def test_adjoint_entry_reports_that_it_is_switched_off():
    """The entry point returns its error before any launch: the mixed form
    is refused as such, the FP64 form with the switch-off message."""

    from wrf_gpu_amd import _lib
    from wrf_gpu_amd.dycore import Dycore, state_to_device

    case, start = _small()
    dyn = Dycore(case.metrics, case.namelist)
    s = state_to_device(start)
    with pytest.raises(_lib.WrfdynError, match="switched off"):
        dyn.step_adjoint(s, {k: v.clone() for k, v in s.items()})
    assert dyn.adjoint_tape_bytes == 0
    dyn.set_precision("mixed")
    with pytest.raises(_lib.WrfdynError, match="FP64"):
        dyn.step_adjoint(s, {k: v.clone() for k, v in s.items()})
    dyn.close()


@pytest.mark.skip(reason=SWITCHED_OFF)
def test_gradient_matches_central_differences_at_257x257x50():
    """The adjoint workload: dJ/dtheta and dJ/du of J = sum(w**2) over the
    top half after 50 steps on the 257x257x50 warm bubble, at the committed
    fixture's points, against its central differences (eps 1e-6)."""

    import torch
    from wrf_gpu_amd import adjoint, ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device

    with open(FIXTURE) as f:
        golden = json.load(f)
    case = ic.bubble_3d(nx=257, ny=257, nz=50)
    nz, n = case.metrics.nz, golden["n_steps"]
    print(f"[adjoint] 257x257x50, {n} steps forward and back", flush=True)
    dyn = Dycore(case.metrics, case.namelist)
    x = {k: v.requires_grad_(k in ("theta", "u")) for k, v in state_to_device(case.state).items()}
    out = adjoint.advance(dyn, x, n)
    J = (out["w"][(nz + 1) // 2:] ** 2).sum()
    assert float(J) == golden["J"]
    J.backward()
    torch.cuda.synchronize()
    worst = 0.0
    for p in golden["points"]:
        g = float(x[p["field"]].grad[tuple(p["index"])])
        fd = p["derivative"]
        err = abs(g - fd) / max(abs(fd), 1e-3)
        worst = max(worst, err)
        print(f"[adjoint] {p['field']:5s} {p['index']} adjoint {g:+.9e} fd {fd:+.9e} rel {err:.2e}", flush=True)
    assert worst <= 1e-4, worst
    dyn.close()
