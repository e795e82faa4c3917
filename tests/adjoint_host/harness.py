"""Build and call the serial host run of the adjoint piece bodies
(adjoint_host.cpp), and compare it with the gradient oracle's piece VJPs.

Test infrastructure only.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "adjoint_host.cpp")
LIB_PATH = os.path.join(HERE, "libadjoint_host.so")
DEPS = [SRC, os.path.join(ROOT, "wrf_gpu_amd", "csrc", "wrfdyn_adjoint.h"),
        os.path.join(ROOT, "wrf_gpu_amd", "csrc", "wrfdyn_common.h"), os.path.join(ROOT, "include", "wrfdyn_math.h")]
_lib = None


def build(force: bool = False) -> str:
    """hipcc, host code only, -ffp-contract=off as the oracle."""

    stale = not os.path.exists(LIB_PATH) or any(os.path.getmtime(d) > os.path.getmtime(LIB_PATH) for d in DEPS)
    if force or stale:
        cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-std=c++17", "-O2", "-fPIC", "-shared",
               "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "wrf_gpu_amd", "csrc"),
               "-I", os.path.join(ROOT, "include"), SRC, "-o", LIB_PATH]
        subprocess.run(cmd, check=True, capture_output=True, text=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.adjh_piece.restype = ctypes.c_int
    return _lib


def host_piece_adjoint(metrics, namelist, arrays: dict, piece_id: int, rk_step: int, dts: float, bar: dict) -> dict:
    """The adjoint bodies' VJP of one piece, run serially on the host."""

    L = lib()
    names = oracle.array_names()
    assert L.adjh_count() == len(names)
    keep = oracle._Keep()
    g = oracle._grid(metrics, keep)
    nl = oracle._namelist(namelist)
    val = {n: np.ascontiguousarray(arrays[n], dtype=np.float64).copy() for n in names}
    out = {n: (np.array(bar[n], dtype=np.float64, copy=True, order="C") if n in bar else np.zeros_like(val[n]))
           for n in names}
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    scr = np.zeros(10 * (nz + 1) * ny * nx)
    rc = L.adjh_piece(ctypes.byref(g), ctypes.byref(nl), ctypes.c_int(piece_id), ctypes.c_int(rk_step),
                      ctypes.c_double(dts),
                      oracle._pointer_table(names, val), oracle._pointer_table(names, out), oracle._ptr(scr))
    if rc != 0:
        raise RuntimeError(f"adjh_piece failed with code {rc}")
    return out


def adjoint_inputs(case, before: dict, piece_id: int, rk_step: int, dts: float) -> dict:
    """The values a piece's adjoint reads: the arrays after the stage setup
    for the setup (its work arrays are written, then read back), the arrays
    before the piece for every other piece."""

    if piece_id != oracle.PIECE_SETUP:
        return before
    after = {k: v.copy() for k, v in before.items()}
    oracle.piece(case.metrics, case.namelist, after, piece_id, rk_step, dts)
    return after


def trace_step(metrics, namelist, state: dict) -> list:
    """[(piece, rk_step, dts, arrays before the piece)] over one large step."""

    arrays = oracle.new_arrays(metrics, state)
    out = []
    for pc in oracle.step_pieces(namelist):
        out.append((*pc, {k: v.copy() for k, v in arrays.items()}))
        oracle.piece(metrics, namelist, arrays, *pc)
    return out
