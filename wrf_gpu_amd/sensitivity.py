"""Sensitivity of a scalar objective of the large step to the initial state.

The objective is the kinetic measure of the upper domain used for the adjoint
workload: ``J = sum(w**2)`` over the w levels of the top half of the column
(levels ``f >= (nz + 1) // 2``) after ``n_steps`` large steps.  Its derivatives
with respect to single initial values of theta or u are taken here by central
finite differences, each from two forward runs of the HIP dycore; they are
the check that a gradient from adjoint kernels has to meet.

All arithmetic runs in the HIP kernels of libwrfdyn.so; torch only holds the
device buffers and sums w**2 in float64.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from .dycore import Dycore, state_to_device


def top_half_w2(state: dict, nz: int) -> float:
    """sum(w**2) over w levels f >= (nz + 1) // 2 of a device state (float64)."""

    w = state["w"]
    return float((w[(nz + 1) // 2:] ** 2).sum())


def objective(case, n_steps: int, state: dict | None = None, device="cuda") -> float:
    """J after ``n_steps`` large steps from ``state`` (default: the case's initial state)."""

    import torch

    dyn = Dycore(case.metrics, case.namelist)
    st = state_to_device(case.state if state is None else state, device)
    dyn.step(st, n_steps)
    torch.cuda.synchronize()
    j = top_half_w2(st, case.metrics.nz)
    dyn.close()
    return j


@dataclasses.dataclass
class FdPoint:
    field: str        # "theta" or "u"
    index: tuple      # (k, j, i) into the field
    eps: float
    j_plus: float
    j_minus: float

    @property
    def derivative(self) -> float:
        return (self.j_plus - self.j_minus) / (2.0 * self.eps)


def seeded_points(case, field: str, count: int, seed: int) -> list:
    """``count`` distinct (k, j, i) indices into ``field``, drawn with ``seed``."""

    shape = case.state[field].shape
    rng = np.random.default_rng(seed)
    flat = rng.choice(int(np.prod(shape)), size=count, replace=False)
    return [tuple(int(v) for v in np.unravel_index(int(f), shape)) for f in flat]


def central_differences(case, n_steps: int, field: str, points, eps: float, device="cuda") -> list:
    """dJ/d(field[p]) ~ (J(+eps) - J(-eps)) / (2 eps) at each point, two forward runs each."""

    import torch

    dyn = Dycore(case.metrics, case.namelist)
    base = state_to_device(case.state, device)
    work = {k: torch.empty_like(v) for k, v in base.items()}
    out = []
    for p in points:
        js = []
        for sign in (1.0, -1.0):
            for k in base:
                work[k].copy_(base[k])
            work[field][p] += sign * eps
            dyn.step(work, n_steps)
            torch.cuda.synchronize()
            js.append(top_half_w2(work, case.metrics.nz))
        out.append(FdPoint(field, tuple(p), eps, js[0], js[1]))
    dyn.close()
    return out
