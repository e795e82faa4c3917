"""WRF-compatible netCDF restart (``wrfrst``) of the dycore carry.

Mirrors the reference's ``write_wrfrst_state`` / ``read_wrfrst_state`` /
``inspect_wrfrst_schema`` (src/gpuwrf/io/wrfrst_netcdf.py:574-698) for the
part of the carry this path advances.  Like the reference, the file has two
layers:

* WRF-named restart variables with WRF dimensions, staggering and attributes
  (``U``, ``V``, ``W``, ``T`` = theta - 300, ``P``, ``PB``, ``PH``, ``PHB``,
  ``MU``, ``MUB``; float32 as WRF writes them; names, dimensions and
  attributes as in src/gpuwrf/io/wrfout_writer.py:447-451,613-700 and
  wrfrst_netcdf.py:367-376), so WRF tools can inspect the state;
* exact float64 extension variables ``GPUWRF_STATE_<FIELD>`` for every state
  field, ``GPUWRF_CARRY_MUDF`` and ``GPUWRF_METRIC_<NAME>`` for the grid
  metrics (the reference's extension naming, wrfrst_netcdf.py:78-79,698-728),
  from which a read resumes bit for bit.

The file is classic netCDF with 64-bit offsets (scipy.io.netcdf_file; the
netCDF4 library the reference uses is not a dependency here).  Reading is
fail-closed on the format, its version and the state field order, and checks
that the WRF layer agrees with the exact layer to float32 rounding.
"""

from __future__ import annotations

import dataclasses
import json
from pathlib import Path

import numpy as np

from .grid import MOIST_FIELDS, STATE_FIELDS, Metrics, Namelist, state_fields
from .restart import _METRIC_ARRAYS, _METRIC_SCALARS, _V1_NAMELIST_DEFAULTS, _host, _json_default

FORMAT = "wrf_gpu_amd.wrfrst"
# 2: a moist state adds WRF's QVAPOR..QGRAUP and their exact layer
FORMAT_VERSION = 2
SUPPORTED_FORMAT_VERSIONS = (1, 2)
STATE_EXTENSION_PREFIX = "GPUWRF_STATE_"
CARRY_EXTENSION_PREFIX = "GPUWRF_CARRY_"
METRIC_EXTENSION_PREFIX = "GPUWRF_METRIC_"
THETA_BASE_OFFSET_K = 300.0

XY = ("Time", "south_north", "west_east")
XYZ = ("Time", "bottom_top", "south_north", "west_east")
U_XYZ = ("Time", "bottom_top", "south_north", "west_east_stag")
V_XYZ = ("Time", "bottom_top", "south_north_stag", "west_east")
W_XYZ = ("Time", "bottom_top_stag", "south_north", "west_east")

# name: (dimensions, MemoryOrder, description, units, stagger, the value from
# the state dict)
_STANDARD = {
    "U": (U_XYZ, "XYZ", "x-wind component", "m s-1", "X", lambda s: s["u"]),
    "V": (V_XYZ, "XYZ", "y-wind component", "m s-1", "Y", lambda s: s["v"]),
    "W": (W_XYZ, "XYZ", "z-wind component", "m s-1", "Z", lambda s: s["w"]),
    "T": (XYZ, "XYZ", "perturbation potential temperature theta-t0", "K", "",
          lambda s: s["theta"] - THETA_BASE_OFFSET_K),
    "P": (XYZ, "XYZ", "perturbation pressure", "Pa", "", lambda s: s["p_perturbation"]),
    "PB": (XYZ, "XYZ", "BASE STATE PRESSURE", "Pa", "", lambda s: s["p_total"] - s["p_perturbation"]),
    "PH": (W_XYZ, "XYZ", "perturbation geopotential", "m2 s-2", "Z", lambda s: s["ph_perturbation"]),
    "PHB": (W_XYZ, "XYZ", "base-state geopotential", "m2 s-2", "Z", lambda s: s["ph_total"] - s["ph_perturbation"]),
    "MU": (XY, "XY ", "perturbation dry air mass in column", "Pa", "", lambda s: s["mu_perturbation"]),
    "MUB": (XY, "XY ", "base state dry air mass in column", "Pa", "", lambda s: s["mu_total"] - s["mu_perturbation"]),
}
# the moisture species under WRF's names (written only for a moist state)
_MOIST_STANDARD = {
    "QVAPOR": ("qv", "Water vapor mixing ratio"),
    "QCLOUD": ("qc", "Cloud water mixing ratio"),
    "QRAIN": ("qr", "Rain water mixing ratio"),
    "QICE": ("qi", "Ice mixing ratio"),
    "QSNOW": ("qs", "Snow mixing ratio"),
    "QGRAUP": ("qg", "Graupel mixing ratio"),
}
_STATE_DIMS = {
    "u": U_XYZ, "v": V_XYZ, "w": W_XYZ, "theta": XYZ, "p_total": XYZ, "p_perturbation": XYZ,
    "ph_total": W_XYZ, "ph_perturbation": W_XYZ, "mu_total": XY, "mu_perturbation": XY, "mudf": XY,
    **{q: XYZ for q in MOIST_FIELDS},
}


def _standard(moist: bool) -> dict:
    """The WRF variables of the standard layer: the dry ones, and QVAPOR..QGRAUP when moist."""

    out = dict(_STANDARD)
    if moist:
        for name, (key, desc) in _MOIST_STANDARD.items():
            out[name] = (XYZ, "XYZ", desc, "kg kg-1", "", lambda s, key=key: s[key])
    return out


def _token(name: str) -> str:
    return "".join(ch.upper() if ch.isalnum() else "_" for ch in name)


def state_extension_name(field: str) -> str:
    return STATE_EXTENSION_PREFIX + _token(field)


def metric_extension_name(field: str) -> str:
    return METRIC_EXTENSION_PREFIX + _token(field)


def _metric_dims(name: str, a: np.ndarray, m: Metrics) -> tuple[str, ...]:
    if a.ndim == 1:
        return ("bottom_top_stag",) if a.shape[0] == m.nz + 1 else ("bottom_top",)
    rows = "south_north_stag" if a.shape[0] == m.ny + 1 else "south_north"
    cols = "west_east_stag" if a.shape[1] == m.nx + 1 else "west_east"
    return (rows, cols)


def write_wrfrst(state: dict, metrics: Metrics, namelist: Namelist, step_index: int, path: str | Path, *,
                 date: str = "0001-01-01_00:00:00") -> Path:
    """Write the carry ``state`` after ``step_index`` large steps as a
    two-layer netCDF restart; returns the path.  Written beside ``path`` and
    renamed over it."""

    from scipy.io import netcdf_file

    fields = state_fields(namelist)
    missing = [k for k in fields if k not in state]
    if missing:
        raise ValueError(f"state lacks {missing}")
    if len(date) != 19:
        raise ValueError("date must be WRF's 19-character YYYY-MM-DD_hh:mm:ss")
    shapes = metrics.shapes()
    host = {}
    for k in fields:
        a = _host(state[k])
        if a.shape != shapes[k]:
            raise ValueError(f"state[{k!r}] has shape {a.shape}, the grid needs {shapes[k]}")
        host[k] = a
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    tmp = target.with_name(target.name + ".tmp")
    m = metrics
    f = netcdf_file(str(tmp), "w", version=2)
    try:
        f.createDimension("Time", 1)
        f.createDimension("DateStrLen", 19)
        for name, n in (("west_east", m.nx), ("south_north", m.ny), ("bottom_top", m.nz),
                        ("west_east_stag", m.nx + 1), ("south_north_stag", m.ny + 1), ("bottom_top_stag", m.nz + 1)):
            f.createDimension(name, n)
        times = f.createVariable("Times", "c", ("Time", "DateStrLen"))
        times[0, :] = np.frombuffer(date.encode(), dtype="S1")
        for name, (dims, order, desc, units, stagger, value) in _standard(namelist.moist).items():
            v = f.createVariable(name, "f4", dims)
            v[0] = np.asarray(value(host), dtype=np.float32)
            v.FieldType = np.int32(104)
            v.MemoryOrder = order
            v.description = desc
            v.units = units
            v.stagger = stagger
        for k in fields:
            name = CARRY_EXTENSION_PREFIX + "MUDF" if k == "mudf" else state_extension_name(k)
            v = f.createVariable(name, "f8", _STATE_DIMS[k])
            v[0] = host[k]
            v.description = f"exact {k} for a bitwise resume"
        for k in _METRIC_ARRAYS:
            a = _host(getattr(m, k))
            v = f.createVariable(metric_extension_name(k), "f8", _metric_dims(k, a, m))
            v[:] = a
        f.TITLE = "wrf_gpu_amd dycore restart"
        f.START_DATE = date
        setattr(f, "WEST-EAST_GRID_DIMENSION", np.int32(m.nx + 1))
        setattr(f, "SOUTH-NORTH_GRID_DIMENSION", np.int32(m.ny + 1))
        setattr(f, "BOTTOM-TOP_GRID_DIMENSION", np.int32(m.nz + 1))
        f.DX = np.float32(m.dx)
        f.DY = np.float32(m.dy)
        f.DT = np.float32(namelist.dt_s)
        f.GPUWRF_FORMAT = FORMAT
        f.GPUWRF_FORMAT_VERSION = np.int32(FORMAT_VERSION)
        f.GPUWRF_STEP_INDEX = np.int32(step_index)
        f.GPUWRF_STATE_FIELD_ORDER = ",".join(fields)
        f.GPUWRF_METRIC_SCALARS = json.dumps({k: getattr(m, k) for k in _METRIC_SCALARS}, default=_json_default)
        f.GPUWRF_NAMELIST = json.dumps(dataclasses.asdict(namelist), default=_json_default)
    finally:
        f.close()
    tmp.replace(target)
    return target


def _attr(f, name: str) -> str:
    v = f._attributes.get(name)
    if v is None:
        raise ValueError(f"not a dycore wrfrst: no {name}")
    return v.decode() if isinstance(v, bytes) else v


def read_wrfrst(path: str | Path, device: str | None = None) -> tuple[dict, Metrics, Namelist, int]:
    """Read a restart from its exact layer: ``(state, metrics, namelist,
    step_index)``, the state as numpy arrays or float64 tensors on
    ``device``."""

    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as f:
        if _attr(f, "GPUWRF_FORMAT") != FORMAT:
            raise ValueError("unsupported wrfrst format")
        version = int(f._attributes.get("GPUWRF_FORMAT_VERSION", -1))
        if version not in SUPPORTED_FORMAT_VERSIONS:
            raise ValueError("unsupported wrfrst version")
        nl = json.loads(_attr(f, "GPUWRF_NAMELIST"))
        if version == 1:
            nl = {**_V1_NAMELIST_DEFAULTS, **nl}
        moist = bool(nl.get("moist", False))
        fields = STATE_FIELDS + (MOIST_FIELDS if moist else ())
        if tuple(_attr(f, "GPUWRF_STATE_FIELD_ORDER").split(",")) != fields:
            raise ValueError("wrfrst state field order does not match the current state schema")
        scal = json.loads(_attr(f, "GPUWRF_METRIC_SCALARS"))
        arrays = {}
        for k in _METRIC_ARRAYS:
            name = metric_extension_name(k)
            if name not in f.variables:
                raise ValueError(f"wrfrst lacks {name}")
            arrays[k] = np.array(f.variables[name][:], dtype=np.float64)
        metrics = Metrics(**{k: scal[k] for k in _METRIC_SCALARS}, **arrays)
        shapes = metrics.shapes()
        state = {}
        for k in fields:
            name = CARRY_EXTENSION_PREFIX + "MUDF" if k == "mudf" else state_extension_name(k)
            if name not in f.variables:
                raise ValueError(f"wrfrst lacks {name}")
            a = np.array(f.variables[name][0], dtype=np.float64)
            if a.shape != shapes[k]:
                raise ValueError(f"wrfrst {name} has shape {a.shape}, the grid needs {shapes[k]}")
            state[k] = np.ascontiguousarray(a)
        for name, (_, _, _, _, _, value) in _standard(moist).items():
            if name not in f.variables:
                raise ValueError(f"wrfrst lacks the WRF variable {name}")
            want = np.asarray(value(state), dtype=np.float32)
            if not np.array_equal(np.array(f.variables[name][0]), want):
                raise ValueError(f"the WRF variable {name} disagrees with the exact layer")
        step = int(f._attributes["GPUWRF_STEP_INDEX"])
    if set(nl) != {fl.name for fl in dataclasses.fields(Namelist)}:
        raise ValueError("wrfrst namelist fields do not match the current namelist schema")
    namelist = Namelist(**nl)
    namelist.validate()
    if device is not None:
        import torch

        state = {k: torch.from_numpy(v).to(device) for k, v in state.items()}
    return state, metrics, namelist, step


def inspect_wrfrst_schema(path: str | Path) -> dict:
    """Variables (dimensions, units) and global attributes, without the data."""

    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as f:
        dec = lambda v: v.decode() if isinstance(v, bytes) else v
        return {
            "variables": {n: {"dimensions": list(v.dimensions), "units": dec(v._attributes.get("units", b""))}
                          for n, v in f.variables.items()},
            "global_attrs": {k: (dec(v) if isinstance(v, (bytes, str)) else np.asarray(v).tolist())
                             for k, v in f._attributes.items()},
        }


__all__ = ["FORMAT", "FORMAT_VERSION", "inspect_wrfrst_schema", "metric_extension_name", "read_wrfrst",
           "state_extension_name", "write_wrfrst"]
