"""Restart of the dycore carry: write the state between large steps and read
it back so that a resumed run continues bit for bit.

Mirrors the reference's ``write_restart`` / ``read_restart`` /
``read_restart_metadata`` (src/gpuwrf/io/restart.py:163-330) for the part of
``OperationalCarry`` this path advances: the prognostic ``State`` fields and
the divergence-damping mass tendency ``mudf`` (operational_mode.py:2064,2254).
Nothing else carries from one large step to the next here: every small-step
work array is rebuilt from the state at each RK stage, so the state and
``mudf`` are the whole carry.  With the grid metrics, the namelist and the
step index, a restart is self-contained.

Like the reference, the file holds host arrays only (the write and the read
are GPU-free), records the field order and a format name and version, and
refuses on read anything whose schema differs (fail-closed).  Unlike the
reference, which pickles its payload, the file is a plain ``.npz`` of float64
arrays plus a JSON header stored as bytes, read with ``allow_pickle=False``,
so loading a restart never executes anything from the file.
"""

from __future__ import annotations

import dataclasses
import json
from pathlib import Path
from typing import Any

import numpy as np

from .grid import STATE_FIELDS, Metrics, Namelist, state_fields

FORMAT = "wrf_gpu_amd.dycore_restart"
# 2: a moist state carries its six species and the namelist its moisture
# options; a version-1 file is a dry state whose namelist predates them
FORMAT_VERSION = 2
SUPPORTED_FORMAT_VERSIONS = (1, 2)
_V1_NAMELIST_DEFAULTS = {"moist": False, "moist_adv_opt": 0}

_METRIC_SCALARS = ("nx", "ny", "nz", "dx", "dy", "p_top", "cf1", "cf2", "cf3")
_METRIC_ARRAYS = tuple(f.name for f in dataclasses.fields(Metrics) if f.name not in _METRIC_SCALARS)


def _json_default(o: Any) -> Any:
    """numpy scalars in the metrics or the namelist as Python numbers."""

    if isinstance(o, np.generic):
        return o.item()
    raise TypeError(f"{type(o).__name__} is not serialisable in a restart header")


def _host(a: Any) -> np.ndarray:
    """A float64 C-contiguous host copy of a numpy array or a torch tensor."""

    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).copy()


def write_restart(state: dict, metrics: Metrics, namelist: Namelist, step_index: int, path: str | Path, *,
                  extra_metadata: dict | None = None) -> Path:
    """Write the carry ``state`` (every field of ``STATE_FIELDS``, and the six
    species when the namelist is moist; numpy arrays or device tensors) after
    ``step_index`` large steps; returns the path.  The file is written beside
    ``path`` and renamed over it, so a reader never sees a partial restart."""

    fields = state_fields(namelist)
    missing = [k for k in fields if k not in state]
    if missing:
        raise ValueError(f"state lacks {missing}")
    shapes = metrics.shapes()
    arrays: dict[str, np.ndarray] = {}
    for k in fields:
        a = _host(state[k])
        if a.shape != shapes[k]:
            raise ValueError(f"state[{k!r}] has shape {a.shape}, the grid needs {shapes[k]}")
        arrays[f"state.{k}"] = a
    for k in _METRIC_ARRAYS:
        arrays[f"metrics.{k}"] = _host(getattr(metrics, k))
    header = {
        "format": FORMAT,
        "format_version": FORMAT_VERSION,
        "state_field_order": list(fields),
        "metric_scalars": {k: getattr(metrics, k) for k in _METRIC_SCALARS},
        "metric_array_order": list(_METRIC_ARRAYS),
        "namelist": dataclasses.asdict(namelist),
        "step_index": int(step_index),
        "metadata": dict(extra_metadata) if extra_metadata else {},
    }
    arrays["header"] = np.frombuffer(json.dumps(header, default=_json_default).encode(), dtype=np.uint8)
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    tmp = target.with_name(target.name + ".tmp")
    with tmp.open("wb") as f:
        np.savez(f, **arrays)
    tmp.replace(target)
    return target


def _read_header(z) -> dict:
    if "header" not in z.files:
        raise ValueError("not a dycore restart: no header")
    header = json.loads(bytes(z["header"]).decode())
    if header.get("format") != FORMAT:
        raise ValueError(f"unsupported restart format {header.get('format')!r}")
    version = int(header.get("format_version", -1))
    if version not in SUPPORTED_FORMAT_VERSIONS:
        raise ValueError(f"unsupported restart version {header.get('format_version')}; "
                         f"supported {SUPPORTED_FORMAT_VERSIONS}")
    nl = header.get("namelist")
    if not isinstance(nl, dict):
        raise ValueError("restart header has no namelist")
    if version == 1:
        nl = {**_V1_NAMELIST_DEFAULTS, **nl}
        header["namelist"] = nl
    fields = STATE_FIELDS + (("qv", "qc", "qr", "qi", "qs", "qg") if nl.get("moist") else ())
    if tuple(header.get("state_field_order", ())) != fields:
        raise ValueError("restart state field order does not match the current state schema")
    if tuple(header.get("metric_array_order", ())) != _METRIC_ARRAYS:
        raise ValueError("restart metric fields do not match the current metrics schema")
    want = {f"state.{k}" for k in fields} | {f"metrics.{k}" for k in _METRIC_ARRAYS} | {"header"}
    if set(z.files) != want:
        raise ValueError("restart arrays do not match its header")
    return header


def read_restart(path: str | Path, device: str | None = None) -> tuple[dict, Metrics, Namelist, int]:
    """Read a restart: ``(state, metrics, namelist, step_index)``.  The state
    fields are numpy arrays, or float64 tensors on ``device`` when given."""

    with np.load(Path(path), allow_pickle=False) as z:
        header = _read_header(z)
        scal = header["metric_scalars"]
        metrics = Metrics(**{k: scal[k] for k in _METRIC_SCALARS},
                          **{k: np.array(z[f"metrics.{k}"]) for k in _METRIC_ARRAYS})
        shapes = metrics.shapes()
        state = {}
        for k in header["state_field_order"]:
            a = np.array(z[f"state.{k}"])
            if a.shape != shapes[k] or a.dtype != np.float64:
                raise ValueError(f"restart state[{k!r}] has shape {a.shape} / {a.dtype}, the grid needs {shapes[k]}")
            state[k] = a
    known = {f.name for f in dataclasses.fields(Namelist)}
    nl = header["namelist"]
    if set(nl) != known:
        raise ValueError("restart namelist fields do not match the current namelist schema")
    namelist = Namelist(**nl)
    namelist.validate()
    if device is not None:
        import torch

        state = {k: torch.from_numpy(v).to(device) for k, v in state.items()}
    return state, metrics, namelist, int(header["step_index"])


def read_restart_metadata(path: str | Path) -> dict:
    """The header of a restart (format, version, step index, grid size, user
    metadata) without reading its arrays."""

    with np.load(Path(path), allow_pickle=False) as z:
        header = _read_header(z)
    scal = header["metric_scalars"]
    return {
        "format": header["format"],
        "format_version": int(header["format_version"]),
        "step_index": int(header["step_index"]),
        "grid": (scal["nz"], scal["ny"], scal["nx"]),
        "state_field_count": len(header["state_field_order"]),
        "metadata": header.get("metadata", {}),
    }


__all__ = ["FORMAT", "FORMAT_VERSION", "SUPPORTED_FORMAT_VERSIONS", "read_restart", "read_restart_metadata",
           "write_restart"]
