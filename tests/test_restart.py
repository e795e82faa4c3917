"""Restart of the dycore carry (wrf_gpu_amd.restart, mirroring
src/gpuwrf/io/restart.py:163-330): the file round-trips every state field,
metric and namelist field exactly; a run resumed from it continues bit for bit
(here on the CPU oracle; test_gpu_restart.py does it on the GPU); and any
schema change is refused on read."""

import json

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic, restart
from wrf_gpu_amd.grid import MOIST_FIELDS, STATE_FIELDS


@pytest.fixture
def case():
    c = ic.stratified_terrain_3d(nx=14, ny=11, nz=6, dx=3000.0, u_sigma=1.0, sigma=0.5)
    nl = c.namelist.replace(diff_6th_opt=2, scalar_adv_opt=1)
    return c.__class__(**{**c.__dict__, "namelist": nl})


def _state(case):
    s = dict(case.state)
    s.setdefault("mudf", np.zeros((case.metrics.ny, case.metrics.nx)))
    return s


def test_round_trip_is_exact(case, tmp_path):
    st = oracle.run(case.metrics, case.namelist, _state(case), 2)
    p = restart.write_restart(st, case.metrics, case.namelist, 2, tmp_path / "r.npz", extra_metadata={"run": "a"})
    got, m, nl, step = restart.read_restart(p)
    assert step == 2 and nl == case.namelist
    for k in STATE_FIELDS:
        assert np.array_equal(got[k], st[k]), k
    for f in ("c1h", "c2h", "c1f", "c2f", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp", "msftx", "msfuy", "msfvx", "eta"):
        assert np.array_equal(getattr(m, f), getattr(case.metrics, f)), f
    assert (m.nx, m.ny, m.nz, m.dx, m.dy, m.cf1, m.cf2, m.cf3, m.p_top) == (
        case.metrics.nx, case.metrics.ny, case.metrics.nz, case.metrics.dx, case.metrics.dy, case.metrics.cf1,
        case.metrics.cf2, case.metrics.cf3, case.metrics.p_top)
    meta = restart.read_restart_metadata(p)
    assert meta["step_index"] == 2 and meta["grid"] == (6, 11, 14) and meta["metadata"] == {"run": "a"}


def test_resumed_run_continues_bit_for_bit(case, tmp_path):
    straight = oracle.run(case.metrics, case.namelist, _state(case), 4)
    half = oracle.run(case.metrics, case.namelist, _state(case), 2)
    p = restart.write_restart(half, case.metrics, case.namelist, 2, tmp_path / "r.npz")
    st, m, nl, step = restart.read_restart(p)
    resumed = oracle.run(m, nl, st, 4 - step)
    for k in STATE_FIELDS:
        assert np.array_equal(resumed[k], straight[k]), k


def _rewrite_header(src, dst, change):
    with np.load(src, allow_pickle=False) as z:
        arrays = {k: np.array(z[k]) for k in z.files}
    header = json.loads(bytes(arrays["header"]).decode())
    change(header, arrays)
    arrays["header"] = np.frombuffer(json.dumps(header).encode(), dtype=np.uint8)
    with open(dst, "wb") as f:
        np.savez(f, **arrays)


@pytest.mark.parametrize("what", ["format", "version", "field_order", "missing_array", "shape", "namelist"])
def test_schema_changes_are_refused(case, tmp_path, what):
    p = restart.write_restart(_state(case), case.metrics, case.namelist, 0, tmp_path / "r.npz")

    def change(h, arrays):
        if what == "format":
            h["format"] = "something else"
        elif what == "version":
            h["format_version"] = 99
        elif what == "field_order":
            h["state_field_order"] = list(reversed(h["state_field_order"]))
        elif what == "missing_array":
            del arrays["state.mudf"]
        elif what == "shape":
            arrays["state.theta"] = arrays["state.theta"][:, :-1]
        else:
            h["namelist"]["not_a_field"] = 1

    bad = tmp_path / "bad.npz"
    _rewrite_header(p, bad, change)
    with pytest.raises(ValueError):
        restart.read_restart(bad)


def test_numpy_scalars_in_the_metrics_and_namelist_are_written(case, tmp_path):
    import dataclasses

    m = dataclasses.replace(case.metrics, nx=np.int64(case.metrics.nx), cf1=np.float32(case.metrics.cf1))
    nl = case.namelist.replace(acoustic_substeps=np.int32(case.namelist.acoustic_substeps))
    p = restart.write_restart(_state(case), m, nl, np.int64(3), tmp_path / "r.npz")
    _, m2, nl2, step = restart.read_restart(p)
    assert step == 3 and m2.nx == case.metrics.nx and nl2.acoustic_substeps == case.namelist.acoustic_substeps
    assert m2.cf1 == float(np.float32(case.metrics.cf1))


def test_write_refuses_a_wrong_state(case, tmp_path):
    st = _state(case)
    st["u"] = st["u"][..., :-1]
    with pytest.raises(ValueError):
        restart.write_restart(st, case.metrics, case.namelist, 0, tmp_path / "r.npz")
    del st["u"]
    with pytest.raises(ValueError):
        restart.write_restart(st, case.metrics, case.namelist, 0, tmp_path / "r.npz")


def test_moist_restart_carries_the_species_and_resumes_bit_for_bit(case, tmp_path):
    mcase = ic.with_moisture(case, moist_adv_opt=2)
    straight = oracle.run(mcase.metrics, mcase.namelist, _state(mcase), 4)
    half = oracle.run(mcase.metrics, mcase.namelist, _state(mcase), 2)
    p = restart.write_restart(half, mcase.metrics, mcase.namelist, 2, tmp_path / "r.npz")
    st, m, nl, step = restart.read_restart(p)
    assert nl.moist and nl.moist_adv_opt == 2
    resumed = oracle.run(m, nl, st, 4 - step)
    for k in STATE_FIELDS + MOIST_FIELDS:
        assert np.array_equal(resumed[k], straight[k]), k
    # a moist namelist needs the species
    dry = {k: v for k, v in half.items() if k not in MOIST_FIELDS}
    with pytest.raises(ValueError):
        restart.write_restart(dry, mcase.metrics, mcase.namelist, 2, tmp_path / "r2.npz")


def test_a_version_1_restart_reads_as_a_dry_state(case, tmp_path):
    p = restart.write_restart(_state(case), case.metrics, case.namelist, 1, tmp_path / "r.npz")

    def as_v1(h, arrays):
        h["format_version"] = 1
        for k in ("moist", "moist_adv_opt"):
            del h["namelist"][k]

    old = tmp_path / "v1.npz"
    _rewrite_header(p, old, as_v1)
    st, _, nl, step = restart.read_restart(old)
    assert step == 1 and nl == case.namelist and not nl.moist
    assert set(st) == set(STATE_FIELDS)
