"""Pin the CPU oracle against the reference's own golden warm-bubble scalars.

The reference holds snapshot scalars of its 2-D Skamarock warm bubble
(80x40, dx=dz=250 m, dt=0.1 s, 10 sound steps) from a JAX-CPU run (v0.13) and
a GPU run (v0.10.0) that agree with each other to ~1e-13
(tests/golden/warm_bubble_2d_reference_scalars.json, extracted from
proofs/v013/cpu_bedrock/f2/skamarock_bubble_diagnostics.json and
proofs/v0100/idealized_baseline_u1/warm_bubble_verdict.json).

Those runs predate two documented dycore changes (cp 1004 -> 7R/2 in
dynamics/acoustic_wrf.py:45-54, the w curvature term in
runtime/operational_mode.py:1784-1822).  The oracle built with exactly those
two constants reverted (liboracle_v013.so) must reproduce the golden
scalars; the current oracle must stay within the drift those changes cause.

The remaining gap (2e-9 to 3e-8 at 100 s) is not rounding: the legacy build with
libm pow in place of the shared one, with the vertical column sums taken from the
top down, or with pg_buoy_w fed the stage-entry EOS pressure instead of the
carried one, moves none of the scalars by more than 1e-12 relative.  It is a
small formula-level difference between the v0.13 code path and the current one
that the oracle follows (v0.14 also reworked advance_w's lid and pg_buoy
handling); the bars below sit just above what is observed.
"""

import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "warm_bubble_2d_reference_scalars.json")) as fh:
        return json.load(fh)["snapshots"]


def _snapshot(case, s):
    """idealized.py:_snapshot (:651-706) diagnostics from a state dict."""

    x = case.info["x_m"]
    z = case.info["z_m"]
    th = s["theta"][:, 0, :] - 300.0
    wf = s["w"][:, 0, :]
    wm = 0.5 * (wf[:-1] + wf[1:])
    uf = s["u"][:, 0, :]
    um = 0.5 * (uf[:, :-1] + uf[:, 1:])
    wp = np.maximum(th, 0.0)
    return {
        "max_abs_w_m_s": float(np.max(np.abs(wm))),
        "theta_prime_max_k": float(np.max(th)),
        "theta_prime_min_k": float(np.min(th)),
        "positive_theta_center_x_m": float(np.sum(wp * x[None, :]) / np.sum(wp)),
        "positive_theta_center_z_m": float(np.sum(wp * z[:, None]) / np.sum(wp)),
        "max_abs_u_m_s": float(np.max(np.abs(um))),
        "mass_total_pa": float(np.sum(s["mu_total"][0])),
        "theta_symmetry_linf_k": float(np.max(np.abs(th - th[:, ::-1]))),
    }


KEYS = ("max_abs_w_m_s", "theta_prime_max_k", "theta_prime_min_k", "positive_theta_center_z_m",
        "max_abs_u_m_s")


def _compare(snap, gold, rtol):
    for k in KEYS:
        rel = abs(snap[k] - gold[k]) / abs(gold[k])
        assert rel <= rtol, (k, snap[k], gold[k], rel)
    assert snap["positive_theta_center_x_m"] == pytest.approx(gold["positive_theta_center_x_m"], rel=1e-12)
    # dry mass is conserved exactly with periodic lateral boundaries
    assert snap["mass_total_pa"] == gold["mass_total_pa"]
    assert snap["theta_symmetry_linf_k"] <= 1e-10


def test_initial_state_matches_reference_ic():
    """The IC restatement reproduces the reference's column dry mass exactly."""

    from wrf_gpu_amd import ic

    case = ic.warm_bubble_2d()
    gold = _golden()["jax_cpu_v013"][0]
    assert float(np.sum(case.state["mu_total"][0])) == gold["mass_total_pa"]


def test_legacy_oracle_reproduces_golden_scalars(oracle_lib):
    """100 s and 250 s snapshots to 5e-8 relative (observed at most 3.0e-8, on theta' min, a value
    of 1e-2 K; 1.5e-9 to 1.5e-8 elsewhere)."""

    from wrf_gpu_amd import ic

    case = ic.warm_bubble_2d()
    gold = _golden()["jax_cpu_v013"]
    s = oracle_lib.run(case.metrics, case.namelist, case.state, 1000, legacy_v013=True)
    _compare(_snapshot(case, s), gold[0], 5e-8)
    s = oracle_lib.run(case.metrics, case.namelist, s, 1500, legacy_v013=True)
    _compare(_snapshot(case, s), gold[1], 5e-8)


def test_legacy_oracle_reproduces_golden_500s(oracle_lib):
    """Full 5000 steps (500 s): 7.5e-7 relative (observed 4.9e-7 on theta' min, at most 8.4e-8
    elsewhere)."""

    from wrf_gpu_amd import ic

    case = ic.warm_bubble_2d()
    gold = _golden()
    s = oracle_lib.run(case.metrics, case.namelist, case.state, 5000, legacy_v013=True)
    snap = _snapshot(case, s)
    _compare(snap, gold["jax_cpu_v013"][2], 7.5e-7)
    _compare(snap, gold["gpu_v0100"][2], 7.5e-7)


def test_current_oracle_within_documented_drift(oracle_lib):
    """The current-constant oracle at 100 s: within 1e-3 of the golden scalars."""

    from wrf_gpu_amd import ic

    case = ic.warm_bubble_2d()
    gold = _golden()["jax_cpu_v013"]
    s = oracle_lib.run(case.metrics, case.namelist, case.state, 1000)
    snap = _snapshot(case, s)
    _compare(snap, gold[0], 1e-3)
    # ...and the constants do matter: the current oracle is NOT the legacy one.
    assert abs(snap["max_abs_w_m_s"] - gold[0]["max_abs_w_m_s"]) > 1e-6
