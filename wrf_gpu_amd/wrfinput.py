"""Ingest of a WRF ``wrfinput``/``wrfout``/``wrfrst`` netCDF file into the
dycore's state and metrics: the step's input side for real cases.

The reference feeds real cases through its ``io`` and ``init/real_init``
packages; this reader covers the part the large step consumes.  From one time
of the file it builds the state in the reference layout
(contracts/state.py:82-111, 811-821) from WRF's perturbation/base pairs:

* ``theta = T + 300`` (``T`` is theta - t0, io/wrfout_writer.py:640-646);
* ``p_total = P + PB``, ``ph_total = PH + PHB``, ``mu_total = MU + MUB``, with
  ``P``, ``PH`` and ``MU`` the perturbations;
* ``u``, ``v``, ``w`` as stored (staggered ``west_east_stag``,
  ``south_north_stag``, ``bottom_top_stag``);
* ``mudf`` zero (a cold start; the carried value only exists in a restart);
* the moisture species ``qv, qc, qr, qi, qs, qg`` from ``QVAPOR``, ``QCLOUD``,
  ``QRAIN``, ``QICE``, ``QSNOW``, ``QGRAUP`` when the file has ``QVAPOR``
  (a species its microphysics does not carry is zero, as in the reference's
  State); step such a state with a moist namelist (``moist=True``).

The metrics come from the vertical coordinate arrays WRF writes (``ZNW``,
``C1H``, ``C2H``, ``C1F``, ``C2F``, ``DNW``, ``RDNW``, ``DN``, ``RDN``,
``FNM``, ``FNP``, ``CF1``, ``CF2``, ``CF3``, ``P_TOP``, and ``C3H``, ``C4H``,
``C3F``, ``C4F`` when present, for hypsometric_opt 2), the six map factors
``MAPFAC_{M,U,V}{X,Y}`` and the ``DX``/``DY`` global attributes, with WRF's
signs (contracts/grid.py:64-164).  The Coriolis parameters ``F``, ``E``,
``SINALPHA`` and ``COSALPHA`` are read as the reference reads them
(dynamics/metrics.py:47-79): a missing ``F`` or ``E`` falls back to
2 Omega sin/cos of ``XLAT`` when the file has it, else to 0, and a missing
``SINALPHA``/``COSALPHA`` to 0 and 1.  Files must be classic netCDF (CDF-1 or
CDF-2), which ``scipy.io.netcdf_file`` reads; netCDF4/HDF5 files need
converting first (``nccopy -k classic``).

The step runs periodic lateral boundaries.  A real-data file (global
attribute ``MAP_PROJ`` nonzero: a map projection, whose lateral boundaries
WRF specifies from ``wrfbdy``) is refused unless the caller asks for it to be
stepped as a periodic domain (``assume_periodic=True``).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from .grid import STATE_FIELDS, Metrics

THETA_BASE_OFFSET_K = 300.0

_VERTICAL = {  # Metrics field: (WRF variable, length is nz + 1)
    "eta": ("ZNW", True), "c1h": ("C1H", False), "c2h": ("C2H", False), "c1f": ("C1F", True),
    "c2f": ("C2F", True), "dnw": ("DNW", False), "rdnw": ("RDNW", False), "dn": ("DN", False),
    "rdn": ("RDN", False), "fnm": ("FNM", False), "fnp": ("FNP", False),
}
_MAPFAC = {"msftx": "MAPFAC_MX", "msfty": "MAPFAC_MY", "msfux": "MAPFAC_UX", "msfuy": "MAPFAC_UY",
           "msfvx": "MAPFAC_VX", "msfvy": "MAPFAC_VY"}
_SCALARS = {"cf1": "CF1", "cf2": "CF2", "cf3": "CF3", "p_top": "P_TOP"}
# Metrics field: (WRF variable, value when absent)
_CORIOLIS = {"f": ("F", 0.0), "e": ("E", 0.0), "sina": ("SINALPHA", 0.0), "cosa": ("COSALPHA", 1.0)}
_MOIST = {"qv": "QVAPOR", "qc": "QCLOUD", "qr": "QRAIN", "qi": "QICE", "qs": "QSNOW", "qg": "QGRAUP"}
# 2 Omega of WRF's EARTH_OMEGA, for F and E from XLAT (dynamics/metrics.py:44)
_TWO_OMEGA = 2.0 * 7.2921e-5


def _var(f, name: str, t: int) -> np.ndarray:
    if name not in f.variables:
        raise ValueError(f"the file lacks the WRF variable {name}")
    v = f.variables[name]
    a = np.array(v[:], dtype=np.float64)
    if v.dimensions and v.dimensions[0] == "Time":
        if not 0 <= t < a.shape[0]:
            raise ValueError(f"time index {t} outside the file's {a.shape[0]} times")
        a = a[t]
    return np.ascontiguousarray(a)


def _gattr(f, name: str) -> float:
    v = f._attributes.get(name)
    if v is None:
        raise ValueError(f"the file lacks the global attribute {name}")
    return float(np.asarray(v).reshape(-1)[0])


def read_wrfinput(path: str | Path, time_index: int = 0, device: str | None = None,
                  assume_periodic: bool = False) -> tuple[dict, Metrics]:
    """``(state, metrics)`` from time ``time_index`` of a WRF netCDF file;
    the state as float64 numpy arrays, or tensors on ``device``.  A real-data
    file (``MAP_PROJ`` nonzero) raises ValueError unless ``assume_periodic``."""

    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as f:
        map_proj = f._attributes.get("MAP_PROJ")
        if map_proj is not None and int(np.asarray(map_proj).reshape(-1)[0]) != 0 and not assume_periodic:
            raise ValueError(
                "the file is a real-data case (MAP_PROJ != 0), whose lateral boundaries WRF specifies from "
                "wrfbdy; this step runs periodic lateral boundaries only (pass assume_periodic=True to step "
                "it as a periodic domain anyway)")
        dims = f.dimensions
        try:
            nx, ny, nz = int(dims["west_east"]), int(dims["south_north"]), int(dims["bottom_top"])
        except KeyError as e:
            raise ValueError(f"the file lacks the WRF dimension {e}") from None
        t = time_index
        vert = {}
        for k, (name, stag) in _VERTICAL.items():
            a = _var(f, name, t).reshape(-1)
            if a.shape[0] != (nz + 1 if stag else nz):
                raise ValueError(f"{name} has {a.shape[0]} levels, the grid has {nz} (+1 staggered)")
            vert[k] = a
        maps = {k: _var(f, name, t) for k, name in _MAPFAC.items()}
        # the dry-pressure coefficients, for hypsometric_opt 2, when the file has them
        for k, (name, stag) in (("c3h", ("C3H", False)), ("c4h", ("C4H", False)), ("c3f", ("C3F", True)),
                                ("c4f", ("C4F", True))):
            if name in f.variables:
                a = _var(f, name, t).reshape(-1)
                if a.shape[0] != (nz + 1 if stag else nz):
                    raise ValueError(f"{name} has {a.shape[0]} levels, the grid has {nz} (+1 staggered)")
                vert[k] = a
        scal = {k: float(_var(f, name, t).reshape(-1)[0]) for k, name in _SCALARS.items()}
        lat = np.deg2rad(_var(f, "XLAT", t)) if "XLAT" in f.variables else None
        rot = {}
        for k, (name, default) in _CORIOLIS.items():
            if name in f.variables:
                a = _var(f, name, t)
            elif k in ("f", "e") and lat is not None:
                a = _TWO_OMEGA * (np.sin(lat) if k == "f" else np.cos(lat))
            else:
                a = np.full((ny, nx), default)
            if a.shape != (ny, nx):
                raise ValueError(f"{name} has shape {a.shape}, the grid needs {(ny, nx)}")
            rot[k] = np.ascontiguousarray(a, dtype=np.float64)
        metrics = Metrics(nx=nx, ny=ny, nz=nz, dx=_gattr(f, "DX"), dy=_gattr(f, "DY"), **scal, **vert, **maps,
                          **rot)
        u, v, w = _var(f, "U", t), _var(f, "V", t), _var(f, "W", t)
        T, P, PB = _var(f, "T", t), _var(f, "P", t), _var(f, "PB", t)
        PH, PHB, MU, MUB = _var(f, "PH", t), _var(f, "PHB", t), _var(f, "MU", t), _var(f, "MUB", t)
        # the moisture species (State qv..qg): with QVAPOR present, all six,
        # a species the file lacks (its microphysics does not carry it) as zeros
        moist = {}
        if "QVAPOR" in f.variables:
            for key, name in _MOIST.items():
                moist[key] = _var(f, name, t) if name in f.variables else np.zeros((nz, ny, nx))
    state = {
        "u": u, "v": v, "w": w, "theta": T + THETA_BASE_OFFSET_K, "p_total": P + PB, "p_perturbation": P,
        "ph_total": PH + PHB, "ph_perturbation": PH, "mu_total": MU + MUB, "mu_perturbation": MU,
        "mudf": np.zeros((ny, nx)), **moist,
    }
    shapes = metrics.shapes()
    for k in state:
        if state[k].shape != shapes[k]:
            raise ValueError(f"{k} has shape {state[k].shape}, the grid needs {shapes[k]}")
    want = {"msftx": (ny, nx), "msfty": (ny, nx), "msfux": (ny, nx + 1), "msfuy": (ny, nx + 1),
            "msfvx": (ny + 1, nx), "msfvy": (ny + 1, nx)}
    for k, shape in want.items():
        if getattr(metrics, k).shape != shape:
            raise ValueError(f"{_MAPFAC[k]} has shape {getattr(metrics, k).shape}, the grid needs {shape}")
    if device is not None:
        import torch

        state = {k: torch.from_numpy(np.ascontiguousarray(a)).to(device) for k, a in state.items()}
    return state, metrics


__all__ = ["read_wrfinput"]
