"""Build and run the host replay of the GPU adjoint's reverse sweep
(adjoint_replay.cpp): the tape is filled from the oracle's traced large step
(the values the GPU's forward recompute stores), the reverse launch sequence
of wd_step_adjoint runs every thread of every block on the CPU, and the
result is the adjoint of the step's starting state.

The executable is built with AddressSanitizer (host code only), with every
tape array its own allocation, so an access outside any tape array ends the
run with a report.  Test infrastructure only.
"""

from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "adjoint_replay.cpp")
EXE = os.path.join(HERE, "adjoint_replay")
DEPS = [SRC, os.path.join(ROOT, "wrf_gpu_amd", "csrc", "wrfdyn_adjoint_glue.h"),
        os.path.join(ROOT, "wrf_gpu_amd", "csrc", "wrfdyn_adjoint.h"),
        os.path.join(ROOT, "wrf_gpu_amd", "csrc", "wrfdyn_common.h"), os.path.join(ROOT, "include", "wrfdyn_math.h")]


def build(force: bool = False) -> str:
    """hipcc, host code only, with AddressSanitizer on the host code."""

    stale = not os.path.exists(EXE) or any(os.path.getmtime(d) > os.path.getmtime(EXE) for d in DEPS)
    if force or stale:
        cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g",
               "-ffp-contract=off", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address",
               "-I", os.path.join(ROOT, "wrf_gpu_amd", "csrc"), "-I", os.path.join(ROOT, "include"), SRC, "-o", EXE]
        subprocess.run(cmd, check=True, capture_output=True, text=True)
    return EXE


def layout() -> dict:
    out = subprocess.run([build(), "layout"], check=True, capture_output=True, text=True).stdout
    rows = dict(line.split(" ", 1) for line in out.strip().splitlines())
    names = oracle.array_names()
    assert int(rows["count"]) == len(names)
    return {k: [names[int(x)] for x in rows[k].split()] for k in ("stage_state", "stage_const", "carried")}


def tape_from_oracle(metrics, namelist, state: dict) -> tuple[list, list, list]:
    """Per stage the state before the setup and the arrays after it, and the
    carried arrays at every boundary (after the setup, then after each
    substep's advance_uv, advance_mu_t and advance_w + calc_p_rho), from the
    oracle's pieces of one large step."""

    arrays = oracle.new_arrays(metrics, state)
    stage_before, stage_after, boundaries = [], [], []
    for piece, rk, dts in oracle.step_pieces(namelist):
        if piece == oracle.PIECE_SETUP:
            stage_before.append({k: v.copy() for k, v in arrays.items()})
        oracle.piece(metrics, namelist, arrays, piece, rk, dts)
        if piece == oracle.PIECE_SETUP:
            stage_after.append({k: v.copy() for k, v in arrays.items()})
        if piece in (oracle.PIECE_SETUP, oracle.PIECE_UV, oracle.PIECE_MU_THETA, oracle.PIECE_P):
            boundaries.append({k: v.copy() for k, v in arrays.items()})
    return stage_before, stage_after, boundaries


def reverse(metrics, namelist, state: dict, bar: dict) -> dict:
    """The replayed reverse sweep: the adjoint of ``state`` (one large step)
    from ``bar``, the adjoint of the state after the step."""

    lay = layout()
    before, after, bnds = tape_from_oracle(metrics, namelist, state)
    m = metrics
    parts = [m.c1h, m.c2h, m.c1f, m.c2f, m.dnw, m.rdnw, m.dn, m.rdn, m.fnm, m.fnp, m.msftx, m.msfty, m.msfux,
             m.msfuy, m.msfvx, m.msfvy]
    for s in range(3):
        parts += [before[s][n] for n in lay["stage_state"]]
        parts += [after[s][n] for n in lay["stage_const"]]
    for b in bnds:
        parts += [b[n] for n in lay["carried"]]
    names = oracle.array_names()
    keys = {"p_pert": "p_perturbation", "ph_pert": "ph_perturbation", "mu_pert": "mu_perturbation"}
    state_names = names[:11]
    parts += [np.asarray(bar[keys.get(n, n)], dtype=np.float64) for n in state_names]
    nl = namelist
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        np.concatenate([np.ascontiguousarray(p, dtype=np.float64).reshape(-1) for p in parts]).tofile(fin)
        args = [build(), "run", fin, fout, str(m.nx), str(m.ny), str(m.nz), str(nl.acoustic_substeps),
                repr(float(m.dx)), repr(float(m.dy)), repr(float(nl.dt_s)), repr(float(nl.epssm)),
                str(int(bool(nl.top_lid))), str(int(nl.w_damping)), str(int(nl.damp_opt)), repr(float(nl.dampcoef)),
                repr(float(nl.zdamp)), repr(float(m.cf1)), repr(float(m.cf2)), repr(float(m.cf3))]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
        r = subprocess.run(args, capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"adjoint_replay failed ({r.returncode}):\n{r.stdout}\n{r.stderr[-4000:]}")
        flat = np.fromfile(fout, dtype=np.float64)
    out, pos = {}, 0
    shapes = metrics.shapes()
    for n in state_names:
        key = keys.get(n, n)
        size = int(np.prod(shapes[key]))
        out[key] = flat[pos:pos + size].reshape(shapes[key])
        pos += size
    assert pos == flat.size
    return out
