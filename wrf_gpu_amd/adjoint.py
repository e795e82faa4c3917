"""Differentiable large steps of the HIP dycore.

``advance(dyn, state, n_steps)`` runs ``n_steps`` large steps and returns the
new state as tensors that torch.autograd can differentiate: the backward pass
is the hand-written adjoint (``wd_step_adjoint``), with one checkpoint of the
state per large step and the step's intermediates recomputed from it on the
GPU.  Every one of the eleven state arrays is an input and an output.

All arithmetic, forward and backward, runs in the HIP kernels of
libwrfdyn.so; torch holds the buffers and the checkpoints.
"""

from __future__ import annotations

from .dycore import _ABI_FIELDS, Dycore

FIELDS = tuple(k for k, _ in _ABI_FIELDS)


def _function():
    import torch

    class _Steps(torch.autograd.Function):
        @staticmethod
        def forward(ctx, dyn: Dycore, n_steps: int, *fields):
            state = {k: f.detach().clone().contiguous() for k, f in zip(FIELDS, fields)}
            checkpoints = []
            for _ in range(n_steps):
                checkpoints.append({k: v.clone() for k, v in state.items()})
                dyn.step(state, 1)
            ctx.dyn = dyn
            ctx.checkpoints = checkpoints
            ctx.like = [f.detach() for f in fields]
            return tuple(state[k] for k in FIELDS)

        @staticmethod
        def backward(ctx, *grads):
            bar = {}
            for k, gr, like in zip(FIELDS, grads, ctx.like):
                bar[k] = torch.zeros_like(like) if gr is None else gr.detach().clone().contiguous()
            for ck in reversed(ctx.checkpoints):
                ctx.dyn.step_adjoint(ck, bar)
            ctx.checkpoints = []
            return (None, None, *[bar[k] for k in FIELDS])

    return _Steps


_STEPS = None


def advance(dyn: Dycore, state: dict, n_steps: int) -> dict:
    """``n_steps`` large steps of ``state`` (float64 device tensors, not
    modified) through the autograd graph; returns the new state dict."""

    global _STEPS
    if _STEPS is None:
        _STEPS = _function()
    if n_steps < 0:
        raise ValueError("n_steps must be >= 0")
    dyn.check_state(state)
    outs = _STEPS.apply(dyn, int(n_steps), *[state[k] for k in FIELDS])
    return dict(zip(FIELDS, outs))
