"""GPU checks of the persistent ensemble step (wd_batch_create_persistent).

One launch per large step runs every kernel body of every member, the
workgroups meeting at a grid barrier between phases.  Each virtual block is
run by the same kernel code, with the block index, grid and thread layout of
the standalone launch, as wd_step, so every member must come out bit for bit
as that member stepped alone: FP64 and mixed, with rotation, at the uneven
ensemble's member sizes, with members tuned to the LDS column tile, over
several launches.  A barrier wait that runs out
(forced here with a zero poll budget) must end every workgroup, be reported
by wd_batch_status and leave the states as they were; the batch then steps
normally.  Members the persistent step cannot run are refused.
"""

import numpy as np
import pytest

from tests.parity_util import FIELDS, field_errors

pytestmark = pytest.mark.gpu


def _assert_bitwise(a: dict, b: dict, what: str):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), (what, k, float(np.max(np.abs(a[k] - b[k]))))


def _small_cases():
    from wrf_gpu_amd import ic

    return [
        ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, u0=5.0),
        ic.stratified_terrain_3d(nx=20, ny=18, nz=12, dx=3000.0, u_sigma=1.0, sigma=0.5),
        ic.bubble_3d(nx=70, ny=19, nz=12, dx=1000.0, u0=10.0),
        ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=2000.0, u_sigma=0.5, sigma=0.2),
        ic.stratified_terrain_3d(nx=9, ny=1, nz=10, dx=2000.0, u_sigma=0.5, sigma=0.2),
    ]


def _rotating(case, lat0):
    """The case on a beta plane (f, e about lat0) with a projection rotation varying in x."""

    import dataclasses

    m = case.metrics
    two_omega = 2.0 * 7.2921e-5
    y = (np.arange(m.ny) + 0.5) * m.dy
    x = (np.arange(m.nx) + 0.5) * m.dx
    lat = np.deg2rad(lat0 + (y[:, None] - y.mean()) / 111e3 + 0.0 * x[None, :])
    alpha = np.deg2rad(6.0) * np.sin(2.0 * np.pi * x[None, :] / (m.nx * m.dx)) + 0.0 * y[:, None]
    m = dataclasses.replace(m, f=two_omega * np.sin(lat), e=two_omega * np.cos(lat), sina=np.sin(alpha),
                            cosa=np.cos(alpha))
    return case.__class__(**{**case.__dict__, "metrics": m})


def _dycore(case, precision, column=None):
    """A handle for the case; ``column`` ("16x16", "march", ...) picks its acoustic column kernel."""

    import os
    from wrf_gpu_amd.dycore import Dycore

    keep = os.environ.pop("WRFDYN_COLUMN_TILE", None)
    if column is not None:
        os.environ["WRFDYN_COLUMN_TILE"] = column
    try:
        return Dycore(case.metrics, case.namelist, precision=precision)
    finally:
        os.environ.pop("WRFDYN_COLUMN_TILE", None)
        if keep is not None:
            os.environ["WRFDYN_COLUMN_TILE"] = keep


def _run_pair(cases, steps, precision="fp64", columns=None, expect_kernels=None, remaps=None):
    """Each case stepped alone and in one persistent ensemble; returns both host states per case.
    ``columns`` gives each member's column kernel choice, ``expect_kernels`` the column kernel
    the persistent step must then run for each (-1 the column march, 0 the 16x16 tile), and
    ``remaps`` each member's workgroup order (1 XCD-contiguous, 0 hardware order)."""

    import torch
    from wrf_gpu_amd.dycore import Ensemble, state_to_device, state_to_host

    columns = columns or [None] * len(cases)
    alone = [(_dycore(c, precision, col), state_to_device(c.state)) for c, col in zip(cases, columns)]
    batched = [(_dycore(c, precision, col), state_to_device(c.state)) for c, col in zip(cases, columns)]
    for m, remap in enumerate(remaps or []):
        for d, _ in (alone[m], batched[m]):
            d.set_tuned_config(d.tuned_config()["column_tile"], remap)
    ens = Ensemble(batched, persistent=True)
    assert ens.branches == 1 and ens.workgroups > 0 and ens.workgroups % 8 == 0
    kernels = [ens.column_kernel(m) for m in range(len(cases))]
    if expect_kernels is not None:
        assert kernels == list(expect_kernels), kernels
    print(f"[persistent] {len(cases)} members, {precision}, {ens.workgroups} workgroups, steps {steps}, "
          f"column kernels {kernels}", flush=True)
    out = []
    done = 0
    for n in steps:
        ens.step(n)
        for d, s in alone:
            d.step(s, n)
        done += n
        ens.status()
        torch.cuda.synchronize()
        out = [(state_to_host(sa), state_to_host(sb)) for (_, sa), (_, sb) in zip(alone, batched)]
        for m, (a, b) in enumerate(out):
            _assert_bitwise(b, a, f"member {m} ({cases[m].name}) after {done} steps")
    ens.close()
    for d, _ in alone + batched:
        d.close()
    return out


def test_persistent_members_match_individual_steps(oracle_lib):
    cases = _small_cases()
    out = _run_pair(cases, (1, 2))
    ref = oracle_lib.run(cases[1].metrics, cases[1].namelist, cases[1].state, 3)
    errs = field_errors(out[1][1], ref)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])


def test_persistent_mixed_members_match_individual_steps():
    _run_pair(_small_cases()[:4], (2, 1), precision="mixed")


def test_persistent_rotating_members_match_individual_steps():
    cases = _small_cases()
    cases = [_rotating(cases[1], 45.0), cases[0], _rotating(cases[3], 30.0)]
    _run_pair(cases, (2,))


def test_persistent_runs_tuned_column_tiles(oracle_lib):
    """Members tuned to the 16x16 LDS column tile run it inside the launch at 1, 3 and 4 levels
    per thread (8, 40, 50 and 63 levels; the damping layer active at 50 and 63), beside a member
    on the column march, one tuned to another tile shape and one too deep for the persistent
    tile (64 levels), which both run as the column march.  Every member comes out bit for bit
    as stepped alone with its own choice; the 50-level tile member also against the oracle."""

    from wrf_gpu_amd import ic

    terrain = dict(dx=3000.0, u_sigma=1.0, sigma=0.5)
    cases = [
        ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, u0=5.0),
        ic.stratified_terrain_3d(nx=20, ny=18, nz=40, **terrain),
        ic.stratified_terrain_3d(nx=20, ny=18, nz=50, **terrain),
        ic.stratified_terrain_3d(nx=21, ny=7, nz=63, **terrain),
        ic.bubble_3d(nx=70, ny=19, nz=12, dx=1000.0, u0=10.0),
        ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=2000.0, u_sigma=0.5, sigma=0.2),
        ic.stratified_terrain_3d(nx=17, ny=9, nz=64, **terrain),
    ]
    columns = ["16x16", "16x16", "16x16", "16x16", "march", "8x16", "16x16"]
    out = _run_pair(cases, (2, 1), columns=columns, expect_kernels=[0, 0, 0, 0, -1, -1, -1])
    ref = oracle_lib.run(cases[2].metrics, cases[2].namelist, cases[2].state, 3)
    errs = field_errors(out[2][1], ref)
    for k in FIELDS:
        assert errs[k] <= 1e-12, (k, errs[k])


def test_persistent_column_tile_with_rotation_and_hardware_order():
    """The column tile inside the launch beside the Coriolis phase (a rotating member), and in
    hardware workgroup order as well as the XCD-contiguous order: bit for bit as stepped alone."""

    from wrf_gpu_amd import ic

    terrain = dict(dx=3000.0, u_sigma=1.0, sigma=0.5)
    cases = [_rotating(ic.stratified_terrain_3d(nx=20, ny=18, nz=30, **terrain), 45.0),
             ic.stratified_terrain_3d(nx=33, ny=17, nz=50, **terrain),
             _rotating(ic.bubble_3d(nx=24, ny=20, nz=10, dx=1000.0, u0=10.0), 30.0)]
    _run_pair(cases, (2,), columns=["16x16", "16x16", "march"], expect_kernels=[0, 0, -1], remaps=[1, 0, 0])


def test_persistent_autotune_keeps_the_faster_column_choice():
    """Ensemble(autotune=True, persistent=True): each member is tuned alone, then the persistent
    step is timed with the members' tuned tiles and with every member on the column march, on
    copies of the states, and the faster is kept.  The members' column kernels in the step follow
    that choice, the states are not touched by the trials, and every member is bit for bit as
    stepped alone."""

    import torch
    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device, state_to_host

    cases = [ic.bubble_3d(nx=64, ny=64, nz=40, dx=1000.0), ic.bubble_3d(nx=48, ny=40, nz=40, dx=1000.0, u0=5.0)]
    alone = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    batched = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    ens = Ensemble(batched, persistent=True, autotune=True, trial_steps=1)
    print(f"[persistent] autotuned {ens.tuned}, trial {ens.persistent_trial}", flush=True)
    for m, (c, (_, s)) in enumerate(zip(cases, batched)):
        _assert_bitwise(state_to_host(s), c.state, f"member {m} after the trials")
    kernels = [ens.column_kernel(m) for m in range(len(cases))]
    trial = ens.persistent_trial
    if trial is None:  # no member picked a tile: nothing to compare
        assert all(cfg["column_tile"] < 0 for cfg in ens.tuned) and kernels == [-1] * len(cases)
    else:
        assert trial["kept"] in ("tuned", "march") and trial["tuned_ms"] > 0 and trial["march_ms"] > 0
        want = [cfg["column_tile"] if trial["kept"] == "tuned" and cfg["column_tile"] == 0 else -1 for cfg in ens.tuned]
        assert kernels == want, (kernels, want)
        for (d, _), cfg in zip(batched, ens.tuned):
            now = d.tuned_config()
            assert now["xcd_remap"] == cfg["xcd_remap"]
            assert now["column_tile"] == (cfg["column_tile"] if trial["kept"] == "tuned" else -1)
    ens.step(2)
    for d, s in alone:
        d.step(s, 2)
    ens.status()
    torch.cuda.synchronize()
    for m, ((_, sa), (_, sb)) in enumerate(zip(alone, batched)):
        _assert_bitwise(state_to_host(sb), state_to_host(sa), f"member {m}")
    ens.close()
    for d, _ in alone + batched:
        d.close()


def test_set_tuned_config_round_trip_and_refusals():
    """wd_set_tuned_config applies a tuning as wd_tuned_config reports it and refuses a shape that
    does not hold the levels, an unknown shape or order, and any tile in the mixed form."""

    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import Dycore

    c = ic.stratified_terrain_3d(nx=20, ny=18, nz=70, dx=3000.0, u_sigma=1.0, sigma=0.5)
    d = Dycore(c.metrics, c.namelist)
    for cfg in ((0, 1), (1, 0), (-1, 1), (-1, 0)):
        d.set_tuned_config(*cfg)
        assert d.tuned_config() == {"column_tile": cfg[0], "xcd_remap": cfg[1]}
    for bad in ((2, 1), (3, 0), (4, 1), (-2, 1), (0, 2)):  # 32x8 and 16x8 hold 64 levels, not 71
        with pytest.raises(_lib.WrfdynError):
            d.set_tuned_config(*bad)
    assert d.tuned_config() == {"column_tile": -1, "xcd_remap": 0}
    d.set_precision("mixed")
    with pytest.raises(_lib.WrfdynError, match="column march"):
        d.set_tuned_config(0, 1)
    d.set_tuned_config(-1, 1)
    d.close()


def test_persistent_mixed_form_keeps_the_column_march():
    """The LDS column tile is FP64 only: a mixed-form member runs the column march."""

    from wrf_gpu_amd import ic

    cases = [ic.stratified_terrain_3d(nx=20, ny=18, nz=40, dx=3000.0, u_sigma=1.0, sigma=0.5)]
    _run_pair(cases, (2,), precision="mixed", columns=["16x16"], expect_kernels=[-1])


def test_persistent_at_the_ensemble_member_sizes():
    """One member of each of the uneven ensemble's sizes (129^2 .. 321^2 columns, 50 levels),
    FP64 and mixed, three large steps."""

    from wrf_gpu_amd import ic

    cases = [ic.ensemble_member(nx, ny, 50, member=8 * q) for q, (nx, ny) in enumerate(ic.ENSEMBLE_SIZES)]
    for precision in ("fp64", "mixed"):
        _run_pair(cases, (3,), precision=precision)


def test_barrier_wait_that_runs_out_is_reported_and_leaves_the_state():
    """With a zero poll budget every workgroup but the last to arrive gives up at the first
    barrier, and the last gives up at the second: every workgroup returns, the later launches
    return at once, the step is reported and the states are untouched (the first two phases
    write only work arrays); with the budget restored the same batch steps to the members'
    results alone."""

    import torch
    from wrf_gpu_amd import _lib
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device, state_to_host

    cases = _small_cases()[:3]
    alone = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    batched = [(Dycore(c.metrics, c.namelist), state_to_device(c.state)) for c in cases]
    ens = Ensemble(batched, persistent=True)
    print("[persistent] forced barrier time-out", flush=True)
    ens.set_spin_limit(0)
    ens.step(2)
    with pytest.raises(_lib.WrfdynError, match="grid barrier"):
        ens.status()
    for m, (c, (_, s)) in enumerate(zip(cases, batched)):
        _assert_bitwise(state_to_host(s), c.state, f"member {m} after the abandoned steps")
    ens.status()  # the flag was cleared
    ens.set_spin_limit(1 << 22)
    ens.step(2)
    for d, s in alone:
        d.step(s, 2)
    ens.status()
    torch.cuda.synchronize()
    for m, ((_, sa), (_, sb)) in enumerate(zip(alone, batched)):
        _assert_bitwise(state_to_host(sb), state_to_host(sa), f"member {m} after recovery")
    ens.close()
    for d, _ in alone + batched:
        d.close()


def test_persistent_refuses_members_it_cannot_run():
    from wrf_gpu_amd import _lib, ic
    from wrf_gpu_amd.dycore import Dycore, Ensemble, state_to_device

    a = ic.bubble_3d(nx=16, ny=12, nz=8, dx=1000.0, u0=5.0)
    filt = a.__class__(**{**a.__dict__, "namelist": a.namelist.replace(diff_6th_opt=2)})
    b = ic.bubble_3d(nx=12, ny=10, nz=8, dx=1000.0, u0=5.0, ns=4)
    da, dfilt, db = (Dycore(c.metrics, c.namelist) for c in (a, filt, b))
    sa, sfilt, sb = (state_to_device(c.state) for c in (a, filt, b))
    with pytest.raises(_lib.WrfdynError, match="diffusion"):
        Ensemble([(da, sa), (dfilt, sfilt)], persistent=True)
    with pytest.raises(_lib.WrfdynError, match="substeps"):
        Ensemble([(da, sa), (db, sb)], persistent=True)
    dm = Dycore(a.metrics, a.namelist, precision="mixed")
    with pytest.raises(_lib.WrfdynError, match="precision"):
        Ensemble([(da, sa), (dm, state_to_device(a.state))], persistent=True)
    for d in (da, dfilt, db, dm):
        d.close()
