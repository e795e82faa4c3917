"""The gradient oracle (oracle/wrfdyn_gradient.cpp, ``oracle.vjp``) against
central finite differences of the double oracle.

The reference holds no gradient to compare with (it has no jax.grad or vjp
anywhere), so the vector-Jacobian product of the oracle's large step is pinned
by the mathematics: for a weight ``bar`` on the final state and a direction
``d`` in one input field, <vjp(bar), d> must equal the derivative of
<bar, step(x + t d)> at t = 0.  The step is piecewise smooth (sign-switched
upwinding, clamps), so a difference quotient near a switch carries an error
of first order in t; each field is therefore checked with fourth-order
central differences at several t, and the best agreement, measured against
the Cauchy-Schwarz scale |grad_f| |d| (so that a small dot product from
cancellation does not inflate it), must reach the bar below.  The correct
gradient reaches at most 7.3e-9 there.  A planted error in one derivative
rule (sin's derivative taken as sin instead of cos, which only the Rayleigh
damping layer uses) lifts every field's best agreement above the bar
(1.6e-8 to 7e-6).
"""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic

# Physical amplitudes of the input directions, per field.
AMPLITUDE = {
    "u": 1.0, "v": 1.0, "w": 0.1, "theta": 0.5, "p_total": 50.0, "p_perturbation": 50.0,
    "ph_total": 50.0, "ph_perturbation": 50.0, "mu_total": 50.0, "mu_perturbation": 50.0,
}
BEST_CS = 1e-8
STEPS = (1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6)


@pytest.fixture(scope="module")
def problem():
    # Terrain, a correlated theta perturbation and zdamp = 10 km so that the
    # Rayleigh damping layer covers the upper faces; two steps in so that every
    # field has structure.
    case = ic.stratified_terrain_3d(nx=12, ny=10, nz=8, dx=3000.0, u_sigma=1.0, sigma=0.5, zdamp=10000.0)
    start = oracle.run(case.metrics, case.namelist, case.state, 2)
    return case, start


def _difference_quotients(case, start, field, d, n, bar, ref, steps):
    def J(t):
        x = dict(start)
        x[field] = start[field] + t * d
        y = oracle.run(case.metrics, case.namelist, x, n)
        return sum(float(np.sum(bar[k] * (y[k] - ref[k]))) for k in bar)

    return [(8.0 * (J(h) - J(-h)) - (J(2 * h) - J(-2 * h))) / (12.0 * h) for h in steps]


def test_recorded_forward_equals_the_oracle(problem):
    case, start = problem
    ref = oracle.run(case.metrics, case.namelist, start, 2)
    _, final = oracle.vjp(case.metrics, case.namelist, start, 2, {})
    for k in ref:
        assert np.array_equal(final[k], ref[k]), k


def test_zero_weights_give_zero_gradient(problem):
    case, start = problem
    grad, _ = oracle.vjp(case.metrics, case.namelist, start, 1, {})
    for k, g in grad.items():
        assert not np.any(g), k


def test_zero_steps_is_the_identity(problem):
    case, start = problem
    rng = np.random.default_rng(3)
    bar = {k: rng.standard_normal(v.shape) for k, v in start.items()}
    grad, final = oracle.vjp(case.metrics, case.namelist, start, 0, bar)
    for k in start:
        assert np.array_equal(grad[k], bar[k]) and np.array_equal(final[k], start[k]), k


@pytest.mark.parametrize("field", list(AMPLITUDE))
def test_directional_derivative_matches_central_differences(problem, field):
    case, start = problem
    n = 2
    ref = oracle.run(case.metrics, case.namelist, start, n)
    rng = np.random.default_rng(0)
    # weights normalised by each field's magnitude, so that no field's
    # rounding swamps the others in the difference quotient
    bar = {k: rng.standard_normal(v.shape) / max(np.abs(v).max(), 1e-3) for k, v in ref.items()}
    grad, _ = oracle.vjp(case.metrics, case.namelist, start, n, bar)
    d = np.random.default_rng(10 + list(AMPLITUDE).index(field)).standard_normal(start[field].shape) * AMPLITUDE[field]
    ad = float(np.sum(grad[field] * d))
    fd = _difference_quotients(case, start, field, d, n, bar, ref, STEPS)
    scale = float(np.linalg.norm(grad[field]) * np.linalg.norm(d))
    err = [abs(q - ad) / scale for q in fd]
    assert min(err) <= BEST_CS, (field, ad, fd, err)


def test_mudf_derivative_through_the_divergence_damping(problem):
    """mudf enters only advance_uv's divergence damping, weakly: it is checked
    with weights on u and v after one step, where its effect is not lost in
    the rounding of the other fields."""

    case, start = problem
    ref = oracle.run(case.metrics, case.namelist, start, 1)
    rng = np.random.default_rng(1)
    bar = {k: rng.standard_normal(ref[k].shape) for k in ("u", "v")}
    grad, _ = oracle.vjp(case.metrics, case.namelist, start, 1, bar)
    d = rng.standard_normal(start["mudf"].shape) * 10.0
    ad = float(np.sum(grad["mudf"] * d))
    fd = _difference_quotients(case, start, "mudf", d, 1, bar, ref, (1e-2, 1e-3))
    assert min(abs(q - ad) / abs(ad) for q in fd) <= 1e-6, (ad, fd)


def test_top_half_w2_sensitivity_to_theta(problem):
    """The adjoint workload's objective, J = sum(w**2) over the w levels
    f >= (nz + 1) // 2, seeded as bar_w = 2 w there: its theta gradient at
    single points against central differences."""

    case, start = problem
    nz, n = case.metrics.nz, 3
    final = oracle.run(case.metrics, case.namelist, start, n)
    bar_w = np.zeros_like(final["w"])
    bar_w[(nz + 1) // 2:] = 2.0 * final["w"][(nz + 1) // 2:]
    grad, _ = oracle.vjp(case.metrics, case.namelist, start, n, {"w": bar_w})

    def J(theta):
        x = dict(start)
        x["theta"] = theta
        w = oracle.run(case.metrics, case.namelist, x, n)["w"]
        return float(np.sum(w[(nz + 1) // 2:] ** 2))

    rng = np.random.default_rng(5)
    for flat in rng.choice(start["theta"].size, size=4, replace=False):
        p = np.unravel_index(int(flat), start["theta"].shape)
        best = np.inf
        for h in (1e-3, 1e-4):
            tp, tm = start["theta"].copy(), start["theta"].copy()
            tp[p] += h
            tm[p] -= h
            best = min(best, abs((J(tp) - J(tm)) / (2 * h) - grad["theta"][p]) / abs(grad["theta"][p]))
        assert best <= 1e-5, (p, best, grad["theta"][p])


def test_pieces_compose_to_the_step_and_their_vjps_to_the_step_vjp(problem):
    """A large step run piece by piece (origin, per stage setup, the four
    substep pieces nsub times, finish) is bit-identical to orc_run, and the
    piece VJPs chained in reverse over that trace give the whole-step VJP.
    Adjoint kernels are checked piece by piece against orc_piece_vjp."""

    case, start = problem
    arrays = oracle.new_arrays(case.metrics, start)
    pieces = oracle.step_pieces(case.namelist)
    assert len(pieces) == 1 + 3 * 2 + 4 * (1 + 3 + 6)
    trace = []
    for pc in pieces:
        trace.append({k: v.copy() for k, v in arrays.items()})
        oracle.piece(case.metrics, case.namelist, arrays, *pc)
    ref = oracle.run(case.metrics, case.namelist, start, 1)
    for key, name in oracle._STATE_ARRAY.items():
        assert np.array_equal(arrays[name], ref[key]), key

    rng = np.random.default_rng(7)
    bar_state = {k: rng.standard_normal(v.shape) for k, v in ref.items()}
    whole, _ = oracle.vjp(case.metrics, case.namelist, start, 1, bar_state)
    bar = {oracle._STATE_ARRAY[k]: v for k, v in bar_state.items()}
    for pc, before in zip(reversed(pieces), reversed(trace)):
        bar = oracle.piece_vjp(case.metrics, case.namelist, before, *pc, bar)
    for key, name in oracle._STATE_ARRAY.items():
        scale = max(float(np.abs(whole[key]).max()), 1e-300)
        assert float(np.abs(bar[name] - whole[key]).max()) <= 1e-12 * scale, key
