"""GPU checks of the graph-captured step, the batched ensemble and the
autotuner, all through the C ABI.

Each form must give bit-identical results to the plain wd_step on the same
inputs (it runs the same kernels on the same buffers), and the plain step is
held to the oracle by test_gpu_parity.py; the oracle is also checked here
once per form at the plain path's 1e-12 bar.
"""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _assert_bitwise(a: dict, b: dict, what: str):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), (what, k, float(np.max(np.abs(a[k] - b[k]))))


def _assert_oracle(got: dict, ref: dict):
    errs = field_errors(got, ref)
    for k in FIELDS:
        assert errs[k] <= TOL, (k, errs[k])


def test_graph_replay_matches_eager_step(oracle_lib):
    """Replays of the captured step, with eager device work and synchronisation
    between them, equal the eager step after every replay."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=24, ny=20, nz=16, dx=3000.0, u_sigma=1.0, sigma=0.5)
    print(f"[graph] case={case.name} grid=24x20x16", flush=True)
    eager, graphed = Dycore(case.metrics, case.namelist), Dycore(case.metrics, case.namelist)
    sa, sb = state_to_device(case.state), state_to_device(case.state)
    graphed.capture_graph(sb)
    torch.cuda.synchronize()
    # capture records only: the state is unchanged until the first replay
    _assert_bitwise(state_to_host(sb), case.state, "after capture")
    done = 0
    for n in (1, 2, 1):
        eager.step(sa, n)
        graphed.graph_step(n)
        done += n
        torch.cuda.synchronize()
        _assert_bitwise(state_to_host(sb), state_to_host(sa), f"replay to step {done}")
        # unrelated eager work between replays
        scratch = torch.randn(1 << 16, device="cuda", dtype=torch.float64)
        float(scratch.sum())
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, done)
    _assert_oracle(state_to_host(sb), ref)
    eager.close()
    graphed.close()


def test_graph_step_without_capture_is_an_error():
    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import Dycore

    case = ic.bubble_3d(nx=8, ny=6, nz=5, dx=1000.0)
    dyn = Dycore(case.metrics, case.namelist)
    lib = _lib.load()
    assert lib.wd_graph_step(dyn._handle, 1, None) == _lib.WD_EINVAL
    assert "capture" in _lib.last_error(dyn._handle)
    dyn.close()


@pytest.mark.parametrize("branches", [1, 2, 4])
def test_batched_members_match_individual_steps(oracle_lib, branches):
    """Members of three different grids advanced in one captured graph equal
    each member stepped alone, bitwise, for 1, 2 and 4 branches."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device, state_to_host

    cases = [
        ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, u0=5.0),
        ic.stratified_terrain_3d(nx=20, ny=18, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5),
        ic.bubble_3d(nx=24, ny=20, nz=10, dx=1000.0, u0=10.0),
        ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=2000.0, u_sigma=0.5, sigma=0.2),
    ]
    print(f"[batch] members={len(cases)} branches={branches}", flush=True)
    alone = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    batched = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    ens = Ensemble(batched, branches=branches)
    assert sorted({ens.branch(m) for m in range(len(cases))}) == list(range(min(branches, len(cases))))
    steps = 0
    for n in (1, 2):
        ens.step(n)
        for d, s in alone:
            d.step(s, n)
        steps += n
        torch.cuda.synchronize()
        for m, ((_, sa), (_, sb)) in enumerate(zip(alone, batched)):
            _assert_bitwise(state_to_host(sb), state_to_host(sa), f"member {m} step {steps}")
    ref = oracle_lib.run(cases[1].metrics, cases[1].namelist, cases[1].state, steps)
    _assert_oracle(state_to_host(batched[1][1]), ref)
    ens.close()
    for d, _ in alone + batched:
        d.close()


def test_autotuned_ensemble_matches_members_alone(oracle_lib):
    """Ensemble(autotune=True) tunes every member alone, times the batch with those picks and
    with every member on the column march, and captures the faster (the LDS-tiled column kernel
    or the column march, either workgroup order); the states are untouched by the trials, and the
    result is still bitwise each member stepped alone, a moist member included."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device, state_to_host

    cases = [
        ic.bubble_3d(nx=24, ny=20, nz=10, dx=1000.0, u0=10.0),
        ic.stratified_terrain_3d(nx=20, ny=18, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5),
        ic.with_moisture(ic.stratified_terrain_3d(nx=16, ny=12, nz=8, dx=2000.0, u_sigma=0.5, sigma=0.2),
                         moist_adv_opt=1),
    ]
    print(f"[batch] autotuned ensemble of {len(cases)} members", flush=True)
    alone = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    batched = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    ens = Ensemble(batched, branches=2, autotune=True, trial_steps=1)
    assert len(ens.tuned) == len(cases) and ens.tuned[2]["column_tile"] == -1  # moist: the column march
    print(f"[batch] picks {ens.tuned}, trial {ens.batch_trial}", flush=True)
    for m, (c, (_, s)) in enumerate(zip(cases, batched)):  # tuning and the batch trial ran on copies
        _assert_bitwise(state_to_host(s), c.state, f"member {m} after tuning")
    trial = ens.batch_trial
    if trial is None:
        assert all(cfg["column_tile"] < 0 for cfg in ens.tuned)
    else:
        assert trial["kept"] in ("tuned", "march") and ens.persistent_trial is None
    for m, cfg in enumerate(ens.tuned):  # the batch runs the kept choice
        want = cfg["column_tile"] if trial is None or trial["kept"] == "tuned" else -1
        assert ens.column_kernel(m) == want and batched[m][0].tuned_config()["column_tile"] == want
    ens.step(2)
    for d, s in alone:
        d.step(s, 2)
    torch.cuda.synchronize()
    for m, ((_, sa), (_, sb)) in enumerate(zip(alone, batched)):
        _assert_bitwise(state_to_host(sb), state_to_host(sa), f"member {m}")
    ref = oracle_lib.run(cases[2].metrics, cases[2].namelist, cases[2].state, 2)
    got = state_to_host(batched[2][1])
    for k in ("qv", "qc", "qr", "qi", "qs", "qg"):
        assert np.array_equal(got[k], ref[k]), k
    ens.close()
    for d, _ in alone + batched:
        d.close()


def test_batch_rejects_duplicate_member():
    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device

    case = ic.bubble_3d(nx=8, ny=6, nz=5, dx=1000.0)
    d = Dycore(case.metrics, case.namelist)
    s1, s2 = state_to_device(case.state), state_to_device(case.state)
    with pytest.raises(_lib.WrfdynError, match="twice"):
        Ensemble([(d, s1), (d, s2)])
    d.close()


def test_autotune_leaves_state_and_results_unchanged(oracle_lib):
    """The autotuner trials on a copy: the caller's state is untouched, and the
    chosen form steps to the same result as the default form."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = ic.stratified_terrain_3d(nx=20, ny=18, nz=50, dx=3000.0, u_sigma=1.0, sigma=0.5)
    print(f"[autotune] case={case.name} grid=20x18x50", flush=True)
    tuned, plain = Dycore(case.metrics, case.namelist), Dycore(case.metrics, case.namelist)
    st, sp = state_to_device(case.state), state_to_device(case.state)
    cfg = tuned.autotune(st, trial_steps=1)
    torch.cuda.synchronize()
    print("[autotune] chose", cfg, flush=True)
    assert cfg["column_tile"] in (-1, 0, 1, 2, 3) and cfg["xcd_remap"] in (0, 1)
    _assert_bitwise(state_to_host(st), case.state, "state after autotune")
    tuned.step(st, 3)
    plain.step(sp, 3)
    torch.cuda.synchronize()
    _assert_bitwise(state_to_host(st), state_to_host(sp), "tuned vs default")
    tuned.close()
    plain.close()

