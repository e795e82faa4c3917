"""Helpers shared by the parity tests (not collected by pytest)."""

import numpy as np

FIELDS = ("u", "v", "w", "theta", "p_total", "p_perturbation", "ph_total", "ph_perturbation",
          "mu_total", "mu_perturbation", "mudf")


def rel_err(a, b):
    """max|a-b| / max|b| (absolute max difference when b is identically zero)."""

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = float(np.max(np.abs(a - b))) if a.size else 0.0
    s = float(np.max(np.abs(b))) if b.size else 0.0
    return d / s if s > 0 else d


def field_errors(got: dict, ref: dict, fields=FIELDS):
    return {k: rel_err(got[k], ref[k]) for k in fields}


def perturbation_errors(got: dict, ref: dict, start: dict):
    """Relative error of the change over the run, per field (a stricter view for
    fields that carry a large constant base such as p_total or theta)."""

    out = {}
    for k in FIELDS:
        d_ref = ref[k] - start[k]
        d_got = got[k] - start[k]
        out[k] = rel_err(d_got, d_ref)
    return out
