"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is the checker for the HIP product path and the timed CPU
baseline.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg import this module; the product never does.

``build()`` compiles ``liboracle.so`` from ``wrfdyn_oracle.c`` with the
Makefile beside it, and ``liboracle_grad.so``: the same source compiled as
C++ with a recording scalar (``wrfdyn_gradient.cpp``), whose ``orc_vjp`` is
the vector-Jacobian product of the oracle's large step (``vjp`` below).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# Variant built with the reference's pre-v0.14 dycore constants (see Makefile).
LEGACY_LIB_PATH = os.path.join(HERE, "liboracle_v013.so")
GRAD_LIB_PATH = os.path.join(HERE, "liboracle_grad.so")

_dp = ctypes.POINTER(ctypes.c_double)


class _Grid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("c1h", _dp), ("c2h", _dp), ("c1f", _dp), ("c2f", _dp),
        ("dnw", _dp), ("rdnw", _dp), ("dn", _dp), ("rdn", _dp),
        ("fnm", _dp), ("fnp", _dp),
        ("cf1", ctypes.c_double), ("cf2", ctypes.c_double), ("cf3", ctypes.c_double),
        ("msftx", _dp), ("msfty", _dp), ("msfux", _dp), ("msfuy", _dp),
        ("msfvx", _dp), ("msfvy", _dp),
        ("c3h", _dp), ("c4h", _dp), ("c3f", _dp), ("c4f", _dp), ("p_top", ctypes.c_double),
        ("f", _dp), ("e", _dp), ("sina", _dp), ("cosa", _dp),
    ]


class _Namelist(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("ns", ctypes.c_int), ("epssm", ctypes.c_double),
        ("top_lid", ctypes.c_int), ("w_damping", ctypes.c_int), ("damp_opt", ctypes.c_int),
        ("dampcoef", ctypes.c_double), ("zdamp", ctypes.c_double),
        ("diff_6th_opt", ctypes.c_int), ("diff_6th_factor", ctypes.c_double),
        ("diff_opt", ctypes.c_int), ("km_opt", ctypes.c_int), ("c_s", ctypes.c_double),
        ("scalar_adv_opt", ctypes.c_int), ("hypsometric_opt", ctypes.c_int),
        ("const_nu_m2_s", ctypes.c_double), ("moist_adv_opt", ctypes.c_int),
    ]


MOIST_SPECIES = ("qv", "qc", "qr", "qi", "qs", "qg")


class _State(ctypes.Structure):
    _fields_ = [(n, _dp) for n in (
        "u", "v", "w", "theta", "p_total", "p_pert", "ph_total", "ph_pert",
        "mu_total", "mu_pert", "mudf") + MOIST_SPECIES]


_STATE_KEYS = {
    "u": "u", "v": "v", "w": "w", "theta": "theta", "p_total": "p_total",
    "p_pert": "p_perturbation", "ph_total": "ph_total", "ph_pert": "ph_perturbation",
    "mu_total": "mu_total", "mu_pert": "mu_perturbation", "mudf": "mudf",
    **{q: q for q in MOIST_SPECIES},
}


def _mk_state(d: dict) -> "_State":
    """The orc_state of a state dict; the moisture pointers NULL when the dict has none."""

    return _State(*[_ptr(d[_STATE_KEYS[n]]) if _STATE_KEYS[n] in d else None for n, _ in _State._fields_])


_libs: dict = {}


def build(force: bool = False) -> str:
    """Compile liboracle.so (gcc, -ffp-contract=off) and return its path."""

    if force or not all(os.path.exists(p) for p in (LIB_PATH, LEGACY_LIB_PATH, GRAD_LIB_PATH)):
        cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True, capture_output=True, text=True)
    return LIB_PATH


def lib(legacy_v013: bool = False) -> ctypes.CDLL:
    path = LEGACY_LIB_PATH if legacy_v013 else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _lib = ctypes.CDLL(path)
        _libs[path] = _lib
        _lib.orc_run.restype = ctypes.c_int
        _lib.orc_run.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(_Namelist),
                                 ctypes.POINTER(_State), ctypes.c_int]
        _lib.orc_calc_coef_w.restype = ctypes.c_int
        _lib.orc_thomas.restype = None
        _lib.orc_alt_from_theta_p.restype = ctypes.c_double
        _lib.orc_alt_from_theta_p.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_p_from_theta_alt.restype = ctypes.c_double
        _lib.orc_p_from_theta_alt.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_advect_scalar.restype = ctypes.c_int
        _lib.orc_surface_w.restype = ctypes.c_int
        _lib.orc_pow.restype = ctypes.c_double
        _lib.orc_pow.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_sin.restype = ctypes.c_double
        _lib.orc_sin.argtypes = [ctypes.c_double]
        _lib.orc_set_threads.restype = None
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
    return _libs[path]


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class _Keep:
    """Holds the contiguous arrays referenced by a ctypes struct."""

    def __init__(self):
        self.arrays = []

    def __call__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.arrays.append(a)
        return _ptr(a)


def _grid(metrics, keep: _Keep) -> _Grid:
    m = metrics
    return _Grid(
        m.nx, m.ny, m.nz, m.dx, m.dy,
        keep(m.c1h), keep(m.c2h), keep(m.c1f), keep(m.c2f), keep(m.dnw), keep(m.rdnw),
        keep(m.dn), keep(m.rdn), keep(m.fnm), keep(m.fnp), m.cf1, m.cf2, m.cf3,
        keep(m.msftx), keep(m.msfty), keep(m.msfux), keep(m.msfuy), keep(m.msfvx), keep(m.msfvy),
        keep(m.c3h), keep(m.c4h), keep(m.c3f), keep(m.c4f), float(m.p_top),
        *((keep(m.f), keep(m.e), keep(m.sina), keep(m.cosa)) if m.rotating else (None,) * 4),
    )


def run(metrics, namelist, state: dict, n_steps: int, threads: int | None = None,
        legacy_v013: bool = False) -> dict:
    """Advance ``n_steps`` large steps on copies of ``state``; returns the new state dict."""

    L = lib(legacy_v013)
    if threads is not None:
        L.orc_set_threads(int(threads))
    keep = _Keep()
    g = _grid(metrics, keep)
    nl = _namelist(namelist)
    out = _state_copy(metrics, state)
    s = _mk_state(out)
    rc = L.orc_run(ctypes.byref(g), ctypes.byref(nl), ctypes.byref(s), int(n_steps))
    if rc != 0:
        raise RuntimeError(f"orc_run failed with code {rc}")
    return out


def _namelist(namelist) -> _Namelist:
    return _Namelist(float(namelist.dt_s), int(namelist.acoustic_substeps), float(namelist.epssm),
                     int(bool(namelist.top_lid)), int(namelist.w_damping), int(namelist.damp_opt),
                     float(namelist.dampcoef), float(namelist.zdamp), int(namelist.diff_6th_opt),
                     float(namelist.diff_6th_factor), int(namelist.diff_opt), int(namelist.km_opt),
                     float(namelist.c_s), int(namelist.scalar_adv_opt), int(namelist.hypsometric_opt),
                     float(namelist.const_nu_m2_s), int(getattr(namelist, "moist_adv_opt", 0)))


def _state_copy(metrics, state: dict) -> dict:
    out = {k: np.array(v, dtype=np.float64, copy=True, order="C") for k, v in state.items()}
    if "mudf" not in out:
        out["mudf"] = np.zeros((metrics.ny, metrics.nx))
    return out


def grad_lib() -> ctypes.CDLL:
    if GRAD_LIB_PATH not in _libs:
        if not os.path.exists(GRAD_LIB_PATH):
            build()
        L = ctypes.CDLL(GRAD_LIB_PATH)
        L.orc_vjp.restype = ctypes.c_int
        L.orc_vjp.argtypes = [ctypes.POINTER(_Grid), ctypes.POINTER(_Namelist), ctypes.POINTER(_State),
                              ctypes.c_int, ctypes.POINTER(_State), ctypes.POINTER(_State),
                              ctypes.POINTER(_State)]
        _libs[GRAD_LIB_PATH] = L
    return _libs[GRAD_LIB_PATH]


def vjp(metrics, namelist, state: dict, n_steps: int, bar: dict) -> tuple:
    """Vector-Jacobian product of ``n_steps`` large steps of the oracle.

    ``bar`` holds the adjoint of each field of the final state (missing keys
    count as zero).  Returns ``(grad, final)``: ``grad`` the adjoint of each
    field of the starting ``state`` (every state array an independent input)
    and ``final`` the state after ``n_steps``, which equals ``run``'s.
    Single-threaded and memory-hungry (one tape node per recorded operation
    of a large step): meant for small grids.
    """

    L = grad_lib()
    keep = _Keep()
    g = _grid(metrics, keep)
    nl = _namelist(namelist)
    start = _state_copy(metrics, state)
    seed = {k: np.ascontiguousarray(bar[k], dtype=np.float64) if k in bar else np.zeros_like(v)
            for k, v in start.items()}
    grad = {k: np.zeros_like(v) for k, v in start.items()}
    final = {k: np.zeros_like(v) for k, v in start.items()}
    mk = _mk_state
    s_in, s_bar, s_grad, s_out = mk(start), mk(seed), mk(grad), mk(final)
    rc = L.orc_vjp(ctypes.byref(g), ctypes.byref(nl), ctypes.byref(s_in), int(n_steps), ctypes.byref(s_bar),
                   ctypes.byref(s_grad), ctypes.byref(s_out))
    if rc != 0:
        raise RuntimeError(f"orc_vjp failed with code {rc}")
    return grad, final


def calc_coef_w(metrics, mut: np.ndarray, c2a: np.ndarray, dts: float, epssm: float, top_lid=True,
                cqw_raw: np.ndarray | None = None):
    """The implicit w operator's (a, alpha, gamma); ``cqw_raw`` (nz+1, ny, nx) the raw
    moisture face factor (None: the dry state)."""

    L = lib()
    keep = _Keep()
    g = _grid(metrics, keep)
    shape = (metrics.nz + 1, metrics.ny, metrics.nx)
    a, alpha, gamma = (np.zeros(shape) for _ in range(3))
    L.orc_calc_coef_w_cq.restype = ctypes.c_int
    L.orc_calc_coef_w_cq(ctypes.byref(g), keep(mut), keep(c2a), ctypes.c_double(dts),
                         ctypes.c_double(epssm), ctypes.c_int(int(top_lid)),
                         None if cqw_raw is None else keep(cqw_raw), _ptr(a), _ptr(alpha), _ptr(gamma))
    return a, alpha, gamma


def thomas(a, alpha, gamma, rhs):
    L = lib()
    n = len(rhs)
    out = np.zeros(n)
    keep = _Keep()
    L.orc_thomas(ctypes.c_int(n), keep(a), keep(alpha), keep(gamma), keep(rhs), _ptr(out))
    return out


def alt_from_theta_p(theta: float, p: float) -> float:
    return float(lib().orc_alt_from_theta_p(theta, p))


def p_from_theta_alt(theta: float, alt: float) -> float:
    return float(lib().orc_p_from_theta_alt(theta, alt))


def advect_limited(metrics, opt: int, u, v, mu_total, q, qold, mu_old, dt: float) -> np.ndarray:
    """The coupled tendency of q (nz, ny, nx) with the positive-definite
    (opt 1) or monotonic (opt 2) limiter (flux_advection.py:736-975)."""

    L = lib()
    L.orc_advect_limited.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, mu_total, q, qold, mu_old)]
    out = np.zeros((metrics.nz, metrics.ny, metrics.nx))
    rc = L.orc_advect_limited(ctypes.byref(g), ctypes.c_int(opt), *[_ptr(a) for a in arrs], ctypes.c_double(dt),
                              _ptr(out))
    if rc != 0:
        raise RuntimeError("orc_advect_limited failed")
    return out


def smag_tendency(metrics, state: dict, c_s: float) -> tuple[dict, np.ndarray]:
    """The coupled tendencies of the 2-D Smagorinsky horizontal diffusion
    (diff_opt 1, km_opt 4; operational_mode.py:2946-2963) of ``state``
    alone, as {"u", "v", "w", "theta"}, and its eddy viscosity xkmh."""

    L = lib()
    L.orc_smag_tendency.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    st = _state_copy(metrics, state)
    s = _mk_state(st)
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    out = {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)), "w": np.zeros((nz + 1, ny, nx)),
           "theta": np.zeros((nz, ny, nx))}
    km = np.zeros((nz, ny, nx))
    rc = L.orc_smag_tendency(ctypes.byref(g), ctypes.c_double(c_s), ctypes.byref(s), _ptr(out["u"]),
                             _ptr(out["v"]), _ptr(out["w"]), _ptr(out["theta"]), _ptr(km))
    if rc != 0:
        raise RuntimeError("orc_smag_tendency failed")
    return out, km


def coriolis_tendency(metrics, state: dict) -> dict:
    """The Coriolis terms of ``state`` alone: the coupled u and v tendencies
    of large_step_coriolis (rk_addtend_dry.py:364-483) as "u", "v", and the
    cosine-Coriolis part of rw_tend (operational_mode.py:1818-1821) as "w"."""

    L = lib()
    L.orc_coriolis_tendency.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    st = _state_copy(metrics, state)
    s = _mk_state(st)
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    out = {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)), "w": np.zeros((nz + 1, ny, nx))}
    rc = L.orc_coriolis_tendency(ctypes.byref(g), ctypes.byref(s), _ptr(out["u"]), _ptr(out["v"]), _ptr(out["w"]))
    if rc != 0:
        raise RuntimeError("orc_coriolis_tendency failed")
    return out


def moist_factors(metrics, state: dict) -> dict:
    """The moisture coupling factors of ``state`` (which holds the six species):
    qtot on mass levels, cqu and cqv on u and v faces
    (acoustic_wrf.py:303-319) and the raw calc_cq cqw on w faces
    (core/advance_w.py:195-219)."""

    L = lib()
    L.orc_moist_factors.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    st = _state_copy(metrics, state)
    s = _mk_state(st)
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    out = {"qtot": np.zeros((nz, ny, nx)), "cqu": np.zeros((nz, ny, nx + 1)), "cqv": np.zeros((nz, ny + 1, nx)),
           "cqw": np.zeros((nz + 1, ny, nx))}
    rc = L.orc_moist_factors(ctypes.byref(g), ctypes.byref(s), _ptr(out["qtot"]), _ptr(out["cqu"]),
                             _ptr(out["cqv"]), _ptr(out["cqw"]))
    if rc != 0:
        raise RuntimeError("orc_moist_factors failed")
    return out


def constk_tendency(metrics, state: dict, nu: float) -> tuple[dict, float]:
    """The coupled tendencies of the constant-K diffusion (const_nu_m2_s = nu;
    operational_mode.py:2878-2929) of ``state`` alone, as {"u", "v", "w",
    "theta"}, and the layer thickness dz it uses."""

    L = lib()
    L.orc_constk_tendency.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    st = _state_copy(metrics, state)
    s = _mk_state(st)
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    out = {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)), "w": np.zeros((nz + 1, ny, nx)),
           "theta": np.zeros((nz, ny, nx))}
    dz = ctypes.c_double(0.0)
    rc = L.orc_constk_tendency(ctypes.byref(g), ctypes.c_double(nu), ctypes.byref(s), _ptr(out["u"]),
                               _ptr(out["v"]), _ptr(out["w"]), _ptr(out["theta"]), ctypes.byref(dz))
    if rc != 0:
        raise RuntimeError("orc_constk_tendency failed")
    return out, dz.value


def sixth_tendency(field: np.ndarray, coef: float) -> np.ndarray:
    """The uncoupled 6th-order filter tendency (explicit_diffusion.py:200-223)
    of a (nlev, nrow, ncol) field, periodic in rows and columns."""

    L = lib()
    L.orc_sixth_tendency.restype = ctypes.c_int
    f = np.ascontiguousarray(field, dtype=np.float64)
    assert f.ndim == 3
    out = np.zeros_like(f)
    rc = L.orc_sixth_tendency(ctypes.c_int(f.shape[0]), ctypes.c_int(f.shape[1]), ctypes.c_int(f.shape[2]),
                              _ptr(f), ctypes.c_double(coef), _ptr(out))
    if rc != 0:
        raise RuntimeError("orc_sixth_tendency failed")
    return out


def det_pow(x: float, y: float) -> float:
    """The deterministic pow shared with the HIP kernels (include/wrfdyn_math.h)."""
    return float(lib().orc_pow(x, y))


def det_sin(x: float) -> float:
    return float(lib().orc_sin(x))


def advect_scalar(metrics, u, v, mu_total, field) -> np.ndarray:
    L = lib()
    keep = _Keep()
    g = _grid(metrics, keep)
    tend = np.zeros((metrics.nz, metrics.ny, metrics.nx))
    rc = L.orc_advect_scalar(ctypes.byref(g), keep(u), keep(v), keep(mu_total), keep(field), _ptr(tend))
    if rc != 0:
        raise RuntimeError("orc_advect_scalar failed")
    return tend


def surface_w(metrics, phb_surface, u, v) -> np.ndarray:
    """Terrain surface w of advance_w for every column (advance_w.py:505-556)."""
    L = lib()
    keep = _Keep()
    g = _grid(metrics, keep)
    out = np.zeros((metrics.ny, metrics.nx))
    rc = L.orc_surface_w(ctypes.byref(g), keep(phb_surface), keep(u), keep(v), _ptr(out))
    if rc != 0:
        raise RuntimeError("orc_surface_w failed")
    return out


# ---------------------------------------------------------------------------
# A large step as a sequence of pieces over named arrays (wrfdyn_oracle.h)
# ---------------------------------------------------------------------------

PIECE_ORIGIN, PIECE_SETUP, PIECE_UV, PIECE_MU_THETA, PIECE_W, PIECE_P, PIECE_FINISH = range(7)
_STATE_ARRAY = {"u": "u", "v": "v", "w": "w", "theta": "theta", "p_total": "p_total",
                "p_perturbation": "p_pert", "ph_total": "ph_total", "ph_perturbation": "ph_pert",
                "mu_total": "mu_total", "mu_perturbation": "mu_pert", "mudf": "mudf"}


def array_names() -> list:
    L = lib()
    L.orc_array_name.restype = ctypes.c_char_p
    return [L.orc_array_name(a).decode() for a in range(L.orc_array_count())]


def _array_shapes(metrics) -> dict:
    nx, ny, nz = metrics.nx, metrics.ny, metrics.nz
    shape = {0: (nz, ny, nx), 1: (nz + 1, ny, nx), 2: (nz, ny, nx + 1), 3: (nz, ny + 1, nx),
             4: (ny, nx), 5: (ny, nx + 1), 6: (ny + 1, nx)}
    L = lib()
    return {name: shape[L.orc_array_extent(a)] for a, name in enumerate(array_names())}


def new_arrays(metrics, state: dict) -> dict:
    """The arrays of a large step: the state's fields (copied) and a zero workspace."""

    arrays = {name: np.zeros(shp) for name, shp in _array_shapes(metrics).items()}
    for key, name in _STATE_ARRAY.items():
        if key in state:
            arrays[name][...] = state[key]
    return arrays


def step_pieces(namelist) -> list:
    """(piece, rk_step, dts) for every piece of one large step, in order."""

    dt, ns = float(namelist.dt_s), int(namelist.acoustic_substeps)
    out = [(PIECE_ORIGIN, 0, 0.0)]
    for rk, dts, nsub in ((1, dt / 3.0, 1), (2, dt / float(ns), max(ns // 2, 1)), (3, dt / float(ns), ns)):
        out.append((PIECE_SETUP, rk, dts))
        for _ in range(nsub):
            out += [(PIECE_UV, rk, dts), (PIECE_MU_THETA, rk, dts), (PIECE_W, rk, dts), (PIECE_P, rk, dts)]
        out.append((PIECE_FINISH, rk, dts))
    return out


def _pointer_table(names, arrays: dict):
    for n in names:
        a = arrays[n]
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], n
    return (_dp * len(names))(*[_ptr(arrays[n]) for n in names])


def piece(metrics, namelist, arrays: dict, piece_id: int, rk_step: int, dts: float) -> None:
    """Run one piece of a large step in place on ``arrays`` (from ``new_arrays``)."""

    L = lib()
    L.orc_piece.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    nl = _namelist(namelist)
    table = _pointer_table(array_names(), arrays)
    rc = L.orc_piece(ctypes.byref(g), ctypes.byref(nl), ctypes.c_int(piece_id), ctypes.c_int(rk_step),
                     ctypes.c_double(dts), table)
    if rc != 0:
        raise RuntimeError(f"orc_piece failed with code {rc}")


def piece_moist(metrics, namelist, arrays: dict, piece_id: int, rk_step: int, dts: float, cqu, cqv, cqwr) -> None:
    """``piece`` for PIECE_UV or PIECE_W with the moisture's coupling factors on the u, v and w faces."""

    L = lib()
    L.orc_piece_moist.restype = ctypes.c_int
    keep = _Keep()
    g = _grid(metrics, keep)
    nl = _namelist(namelist)
    table = _pointer_table(array_names(), arrays)
    factors = [np.ascontiguousarray(x, dtype=np.float64) for x in (cqu, cqv, cqwr)]
    rc = L.orc_piece_moist(ctypes.byref(g), ctypes.byref(nl), ctypes.c_int(piece_id), ctypes.c_int(rk_step),
                           ctypes.c_double(dts), table, *[_ptr(x) for x in factors])
    if rc != 0:
        raise RuntimeError(f"orc_piece_moist failed with code {rc}")


def piece_vjp(metrics, namelist, arrays: dict, piece_id: int, rk_step: int, dts: float, bar: dict) -> dict:
    """Adjoint of every array before one piece, from ``bar`` (the adjoint of
    the arrays after it; missing names count as zero).  ``arrays`` holds the
    values before the piece and is not modified."""

    L = grad_lib()
    L.orc_piece_vjp.restype = ctypes.c_int
    names = array_names()
    keep = _Keep()
    g = _grid(metrics, keep)
    nl = _namelist(namelist)
    before = {n: np.ascontiguousarray(arrays[n], dtype=np.float64).copy() for n in names}
    out = {n: (np.array(bar[n], dtype=np.float64, copy=True, order="C") if n in bar else np.zeros_like(before[n]))
           for n in names}
    rc = L.orc_piece_vjp(ctypes.byref(g), ctypes.byref(nl), ctypes.c_int(piece_id), ctypes.c_int(rk_step),
                         ctypes.c_double(dts), _pointer_table(names, before), _pointer_table(names, out))
    if rc != 0:
        raise RuntimeError(f"orc_piece_vjp failed with code {rc}")
    return out
