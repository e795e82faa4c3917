"""The adjoint bodies of every piece of the large step
(wrf_gpu_amd/csrc/wrfdyn_adjoint.h), run serially on the host, against the
gradient oracle's per-piece VJPs (oracle.piece_vjp).

Each piece's inputs are taken from a real large step of the oracle (piece by
piece, tests/adjoint_host/harness.trace_step), and a random bar on every
array after the piece must map to the oracle's bar before it to 1e-12 of
each array's largest entry (the two differ only in the order in which
contributions are summed).  The stage setup's adjoint reads the arrays as
they are after the setup (it writes only work arrays, which it reads back);
every other piece reads them as they were before it.  Chained in reverse over
a whole large step, the pieces give the gradient oracle's step VJP.
"""

import numpy as np
import pytest

from oracle import oracle
from tests.adjoint_host import harness
from wrf_gpu_amd import ic

PIECES = {"setup": oracle.PIECE_SETUP, "uv": oracle.PIECE_UV, "mu_theta": oracle.PIECE_MU_THETA,
          "w": oracle.PIECE_W, "p": oracle.PIECE_P, "finish": oracle.PIECE_FINISH}
TOL = 1e-12


@pytest.fixture(scope="module")
def traced():
    # terrain (surface w, varying map-free metrics), Rayleigh damping over the
    # upper faces (zdamp = 10 km), two steps in so every field has structure
    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=3000.0, u_sigma=1.0, sigma=0.5, zdamp=10000.0)
    start = oracle.run(case.metrics, case.namelist, case.state, 2)
    return case, harness.trace_step(case.metrics, case.namelist, start)


def _compare(case, before, piece_id, rk, dts, seed):
    rng = np.random.default_rng(seed)
    bar = {k: rng.standard_normal(v.shape) for k, v in before.items()}
    ref = oracle.piece_vjp(case.metrics, case.namelist, before, piece_id, rk, dts, bar)
    got = harness.host_piece_adjoint(case.metrics, case.namelist, harness.adjoint_inputs(case, before, piece_id, rk, dts),
                                     piece_id, rk, dts, bar)
    changed = 0
    for k in ref:
        scale = max(float(np.abs(ref[k]).max()), 1e-300)
        err = float(np.abs(got[k] - ref[k]).max()) / scale
        assert err <= TOL, (k, err)
        changed += not np.array_equal(ref[k], bar[k])
    assert changed >= 8  # the piece moved bars into several arrays


@pytest.mark.parametrize("rk", [1, 2, 3])
@pytest.mark.parametrize("name", list(PIECES))
def test_substep_piece_adjoint_matches_oracle_vjp(traced, name, rk):
    case, trace = traced
    piece_id = PIECES[name]
    steps = [t for t in trace if t[0] == piece_id and t[1] == rk]
    pc, rk_, dts, before = steps[-1]
    _compare(case, before, pc, rk_, dts, seed=10 * piece_id + rk)


def test_advance_w_adjoint_with_w_damping_active(traced):
    """Vertical Courant numbers above 1 switch w_damping's term on; strong
    omega and w are planted into the inputs so that about a fifth of the
    interior faces take it."""

    case, trace = traced
    pc, rk, dts, before = [t for t in trace if t[0] == oracle.PIECE_W][-1]
    before = dict(before)
    rng = np.random.default_rng(1)
    before["ww"] = before["ww"] + rng.standard_normal(before["ww"].shape) * 3000.0
    before["w_work"] = before["w_work"] + rng.standard_normal(before["w_work"].shape) * 10.0
    m, nz = case.metrics, case.metrics.nz
    above = sum(int(np.sum(np.abs(before["ww"][f] / (m.c1f[f] * before["mut"] + m.c2f[f]) * m.rdnw[f] * dts) > 1.0))
                for f in range(1, nz))
    assert above > 50
    _compare(case, before, pc, rk, dts, seed=99)


def test_origin_adjoint_matches_oracle_vjp(traced):
    case, trace = traced
    pc, rk, dts, before = trace[0]
    assert pc == oracle.PIECE_ORIGIN
    rng = np.random.default_rng(3)
    bar = {k: rng.standard_normal(v.shape) for k, v in before.items()}
    ref = oracle.piece_vjp(case.metrics, case.namelist, before, pc, rk, dts, bar)
    got = harness.host_piece_adjoint(case.metrics, case.namelist, before, pc, rk, dts, bar)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_chained_piece_adjoints_give_the_step_vjp(traced):
    """Every piece's adjoint, chained from the last piece of a large step to
    the first, against orc_vjp for that step."""

    case, trace = traced
    start = {key: trace[0][3][name] for key, name in oracle._STATE_ARRAY.items()}
    final = oracle.run(case.metrics, case.namelist, start, 1)
    rng = np.random.default_rng(11)
    bar_state = {k: rng.standard_normal(v.shape) for k, v in final.items()}
    whole, _ = oracle.vjp(case.metrics, case.namelist, start, 1, bar_state)
    bar = {oracle._STATE_ARRAY[k]: v for k, v in bar_state.items()}
    for pc, rk, dts, before in reversed(trace):
        bar = harness.host_piece_adjoint(case.metrics, case.namelist,
                                         harness.adjoint_inputs(case, before, pc, rk, dts), pc, rk, dts, bar)
    for key, name in oracle._STATE_ARRAY.items():
        scale = max(float(np.abs(whole[key]).max()), 1e-300)
        assert float(np.abs(bar[name] - whole[key]).max()) <= 1e-11 * scale, key
