"""GPU parity of the moist large step against the CPU oracle, bit for bit in FP64: the six
species couple into the dynamics (cqu/cqv in the large-step and acoustic PGFs, the moist
rw_tend and implicit-w cqw; dynamics/acoustic_wrf.py:303-319, core/advance_w.py:195-290,
core/rk_addtend_dry.py:278, :359-360) and, with moist_adv_opt 1 or 2, are transported
(runtime/operational_mode.py:3162-3316).  Covered: terrain with each moist_adv_opt, the
2-D slab, the theta limiter and the species' limiter sharing their scratch on the final
stage, the diffusions, a captured graph, wd_step_noalloc, and the autotuner (which must
leave the caller's species untouched and keep the column march); every option at once
(rotation, hypsometric_opt 2, both limiters, the filter); a state whose species are zero
steps as the dry state; and the handle refuses mismatched states."""

import ctypes

import numpy as np
import pytest

from tests.parity_util import FIELDS

pytestmark = pytest.mark.gpu

SPECIES = ("qv", "qc", "qr", "qi", "qs", "qg")


def _with(case, **kw):
    return case.__class__(**{**case.__dict__, "namelist": case.namelist.replace(**kw)})


def _moist_case(nx, ny, nz, opt, **kw):
    from wrf_gpu_amd import ic

    base = ic.stratified_terrain_3d(nx=nx, ny=ny, nz=nz, dx=3000.0, u_sigma=1.0, sigma=0.5)
    return _with(ic.with_moisture(base, moist_adv_opt=opt), **kw)


def _gpu(case, n_steps, mode="step"):
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host, step_noalloc

    m, nl = case.metrics, case.namelist
    print(f"[moist] case={case.name} grid={m.nx}x{m.ny}x{m.nz} steps={n_steps} mode={mode} "
          f"moist_adv_opt={nl.moist_adv_opt} scalar_adv_opt={nl.scalar_adv_opt}", flush=True)
    st = state_to_device(case.state, "cuda")
    if mode == "noalloc":
        step_noalloc(st, nl, m, n_steps)
    else:
        dyn = Dycore(m, nl)
        if mode == "graph":
            dyn.capture_graph(st)
            dyn.graph_step(n_steps)
        else:
            dyn.step(st, n_steps)
        torch.cuda.synchronize()
        dyn.close()
    torch.cuda.synchronize()
    return state_to_host(st)


def _rotating(m, lat0=45.0):
    """A beta-plane f and e = 2 Omega cos(lat) about lat0, and a projection rotation varying in x."""

    import dataclasses

    two_omega = 2.0 * 7.2921e-5
    y = (np.arange(m.ny) + 0.5) * m.dy
    x = (np.arange(m.nx) + 0.5) * m.dx
    lat = np.deg2rad(lat0 + (y[:, None] - y.mean()) / 111e3 + 0.0 * x[None, :])
    alpha = np.deg2rad(8.0) * np.sin(2.0 * np.pi * x[None, :] / (m.nx * m.dx)) + 0.0 * y[:, None]
    return dataclasses.replace(m, f=two_omega * np.sin(lat), e=two_omega * np.cos(lat), sina=np.sin(alpha),
                               cosa=np.cos(alpha))


def _check_bitwise(got, ref, keys):
    diff = {k: float(np.max(np.abs(got[k] - ref[k]))) for k in keys}
    print("max |gpu - oracle|:", {k: f"{v:.1e}" for k, v in diff.items()}, flush=True)
    for k in keys:
        assert np.array_equal(got[k], ref[k]), (k, diff[k])


@pytest.mark.parametrize("opt", [0, 1, 2])
def test_moist_terrain_matches_oracle_bitwise(oracle_lib, opt):
    case = _moist_case(24, 20, 12, opt)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 3)
    got = _gpu(case, 3)
    _check_bitwise(got, ref, FIELDS + SPECIES)
    if opt:
        assert not np.array_equal(ref["qc"], case.state["qc"])  # the species moved


def test_moist_slab_with_both_limiters_and_diffusions(oracle_lib):
    from wrf_gpu_amd import ic

    base = ic.warm_bubble_2d()
    case = _with(ic.with_moisture(base, moist_adv_opt=2), scalar_adv_opt=1, diff_6th_opt=2)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 6)
    got = _gpu(case, 6)
    _check_bitwise(got, ref, FIELDS + SPECIES)


def test_moist_with_smagorinsky_graph_replay(oracle_lib):
    case = _moist_case(20, 16, 10, 1, diff_opt=1, km_opt=4, scalar_adv_opt=2)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 3)
    got = _gpu(case, 3, mode="graph")
    _check_bitwise(got, ref, FIELDS + SPECIES)


def test_moist_noalloc_matches_oracle(oracle_lib):
    case = _moist_case(16, 12, 8, 1)
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 2)
    got = _gpu(case, 2, mode="noalloc")
    _check_bitwise(got, ref, FIELDS + SPECIES)


def test_zero_moisture_is_the_dry_step_on_the_gpu():
    from wrf_gpu_amd import ic

    dry = _with(ic.stratified_terrain_3d(nx=20, ny=16, nz=10, dx=3000.0, u_sigma=1.0, sigma=0.5))
    state = dict(dry.state)
    for q in SPECIES:
        state[q] = np.zeros_like(state["theta"])
    moist = dry.__class__(**{**dry.__dict__, "state": state,
                             "namelist": dry.namelist.replace(moist=True, moist_adv_opt=2)})
    a = _gpu(dry, 2)
    b = _gpu(moist, 2)
    _check_bitwise(b, a, FIELDS)
    for q in SPECIES:
        assert not np.any(b[q])


def test_autotune_keeps_species_and_the_column_march():
    import torch
    from wrf_gpu_amd.dycore import Dycore, state_to_device, state_to_host

    case = _moist_case(16, 12, 8, 2)
    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state, "cuda")
    cfg = dyn.autotune(st, trial_steps=1)
    torch.cuda.synchronize()
    after = state_to_host(st)
    dyn.close()
    assert cfg["column_tile"] == -1
    for k in FIELDS + SPECIES:
        assert np.array_equal(after[k], case.state[k]), k


def test_handles_refuse_mismatched_states():
    from wrf_gpu_amd import _lib
    from wrf_gpu_amd.dycore import Dycore, _state_ptrs, state_to_device

    case = _moist_case(16, 12, 8, 1)
    st = state_to_device(case.state, "cuda")
    moist = Dycore(case.metrics, case.namelist)
    dry = Dycore(case.metrics, case.namelist.replace(moist=False, moist_adv_opt=0))
    lib = _lib.load()
    # a moist handle given a state without species, a dry one given species
    code = lib.wd_step(moist._handle, ctypes.byref(_state_ptrs(st, False)), 1, None)
    assert code == _lib.WD_EINVAL and "six moisture species" in _lib.last_error(moist._handle)
    code = lib.wd_step(dry._handle, ctypes.byref(_state_ptrs(st, True)), 1, None)
    assert code == _lib.WD_EINVAL and "moist = 0" in _lib.last_error(dry._handle)
    with pytest.raises(_lib.WrfdynError):
        moist.set_precision("mixed")
    moist.close()
    dry.close()


@pytest.mark.parametrize("grid", [(40, 36, 16), (21, 7, 9)])
def test_moist_with_rotation_log_pressure_and_filter(oracle_lib, grid):
    """Every option at once: moisture with the monotonic species limiter, a rotating grid
    (beta-plane f, e, a varying projection rotation), hypsometric_opt 2 with c3 = eta,
    the positive-definite theta limiter and the 6th-order filter."""

    import dataclasses

    nx, ny, nz = grid
    case = _moist_case(nx, ny, nz, 2, hypsometric_opt=2, scalar_adv_opt=1, diff_6th_opt=2)
    m = case.metrics
    eta_mass = 0.5 * (m.eta[:-1] + m.eta[1:])
    m = _rotating(dataclasses.replace(m, c3h=eta_mass, c4h=np.zeros(nz), c3f=m.eta.copy(), c4f=np.zeros(nz + 1)))
    case = case.__class__(**{**case.__dict__, "metrics": m})
    ref = oracle_lib.run(case.metrics, case.namelist, case.state, 3)
    got = _gpu(case, 3)
    _check_bitwise(got, ref, FIELDS + SPECIES)
