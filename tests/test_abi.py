"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/wrfdyn.h declares, the ctypes structs match the C layout, and the
argument checks that run before any device call return the documented codes.
"""

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from wrf_gpu_amd import _lib
from wrf_gpu_amd.grid import Metrics, Namelist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wrfdyn.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(wd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert set(declared) == set(_lib.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def _c_layout(structs):
    """sizeof and every field offset of the named C structs, from the header."""

    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "wrfdyn.h"', "int main(void) {"]
    for cname, fields in structs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def _py_layout(structs):
    out = {}
    for cname, cls in structs.items():
        out[cname] = ctypes.sizeof(cls)
        for f, _ in cls._fields_:
            out[f"{cname}.{f}"] = getattr(cls, f).offset
    return out


def _header_fields(cname):
    """The member names of a typedef struct in include/wrfdyn.h, in order."""

    text = open(HEADER).read()
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + cname + ";", text).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            names.append(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", part)[-1])
    return names


def test_ctypes_structs_match_c_layout():
    """_lib.py's structs have the header's members, in order, at the header's offsets."""

    py = {"wd_grid": _lib.WdGrid, "wd_namelist": _lib.WdNamelist, "wd_state": _lib.WdState}
    for cname, cls in py.items():
        assert [f for f, _ in cls._fields_] == _header_fields(cname), cname
    c = _c_layout({k: [f for f, _ in v._fields_] for k, v in py.items()})
    assert _py_layout(py) == c


def _documented_binding():
    """The ctypes structs of the reference-side binding shown in INTEGRATION.md."""

    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    code = re.search(r"```python\n(.*?)```", text, re.S).group(1)
    defs = code[code.index("_dp = "):code.index("_lib.wd_create.argtypes")]
    ns = {"ctypes": ctypes}
    exec(defs, ns)
    return code, {k: ns[k] for k in ("wd_grid", "wd_namelist", "wd_state")}


def test_documented_binding_matches_c_layout():
    """A binding copied from INTEGRATION.md has the header's members and layout, and its
    create() fills every namelist member."""

    code, doc = _documented_binding()
    for cname, cls in doc.items():
        assert [f for f, _ in cls._fields_] == _header_fields(cname), cname
    assert _py_layout(doc) == _c_layout({k: [f for f, _ in v._fields_] for k, v in doc.items()})
    call = re.search(r"nl = wd_namelist\((.*?)\)\n", code, re.S).group(1)
    assert set(re.findall(r"(\w+)=", call)) == {f for f, _ in doc["wd_namelist"]._fields_}


def test_create_rejects_bad_arguments_without_touching_the_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.wd_create(None, None, ctypes.byref(h)) == _lib.WD_EINVAL
    assert "NULL" in _lib.last_error(None)
    grid = _lib.WdGrid()
    grid.nx, grid.ny, grid.nz = 3, 1, 10  # nx < 4
    nl = _lib.WdNamelist()
    assert lib.wd_create(ctypes.byref(grid), ctypes.byref(nl), ctypes.byref(h)) == _lib.WD_EINVAL
    assert "grid too small" in _lib.last_error(None)
    assert lib.wd_step(None, None, 1, None) == _lib.WD_EINVAL
    assert lib.wd_workspace_bytes(None) == 0
    lib.wd_destroy(None)


def test_create_validates_the_moisture_options_on_the_host():
    """moist must be 0 or 1, moist_adv_opt 0..2 and only with moist: refused before any
    device call (here without a GPU)."""

    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import _abi_structs

    lib = _lib.load()
    case = ic.bubble_3d(nx=8, ny=6, nz=6, noise=0.0)
    keep: list = []
    grid, nl = _abi_structs(case.metrics, case.namelist, keep)
    h = ctypes.c_void_p()
    for moist, opt, msg in ((2, 0, "moist must be"), (1, 3, "moist_adv_opt must be"), (0, 1, "needs moist")):
        nl.moist, nl.moist_adv_opt = moist, opt
        assert lib.wd_create(ctypes.byref(grid), ctypes.byref(nl), ctypes.byref(h)) == _lib.WD_EINVAL
        assert msg in _lib.last_error(None)


def test_graph_batch_autotune_and_adjoint_entry_points_reject_bad_arguments():
    lib = _lib.load()
    assert lib.wd_capture_graph(None, None) == _lib.WD_EINVAL
    assert lib.wd_graph_step(None, 1, None) == _lib.WD_EINVAL
    assert lib.wd_autotune(None, None, 1, None) == _lib.WD_EINVAL
    t, r = ctypes.c_int(), ctypes.c_int()
    assert lib.wd_tuned_config(None, ctypes.byref(t), ctypes.byref(r)) == _lib.WD_EINVAL
    assert lib.wd_set_precision(None, _lib.WD_PREC_MIXED) == _lib.WD_EINVAL
    assert lib.wd_get_precision(None) == _lib.WD_EINVAL
    b = ctypes.c_void_p()
    assert lib.wd_batch_create(None, None, 0, 1, ctypes.byref(b)) == _lib.WD_EINVAL
    assert "member" in _lib.last_error(None)
    handles = (ctypes.c_void_p * 1)(None)
    states = (_lib.WdState * 1)()
    assert lib.wd_batch_create(handles, states, 1, 0, ctypes.byref(b)) == _lib.WD_EINVAL
    assert "branches" in _lib.last_error(None)
    assert lib.wd_step_batched(None, 1, None) == _lib.WD_EINVAL
    assert lib.wd_batch_branch(None, 0) == _lib.WD_EINVAL
    assert lib.wd_batch_branches(None) == _lib.WD_EINVAL
    assert lib.wd_batch_last_error(None) is not None
    lib.wd_batch_destroy(None)
    assert lib.wd_step_adjoint(None, None, None, None) == _lib.WD_EINVAL
    assert lib.wd_adjoint_tape_bytes(None) == 0
    # the persistent batch: refused before any device call
    assert lib.wd_batch_create_persistent(None, None, 0, ctypes.byref(b)) == _lib.WD_EINVAL
    assert "member" in _lib.last_error(None)
    assert lib.wd_batch_create_persistent(handles, states, 1, ctypes.byref(b)) == _lib.WD_EINVAL
    assert "NULL" in _lib.last_error(None)
    assert lib.wd_batch_create_persistent(handles, states, 1, None) == _lib.WD_EINVAL
    assert lib.wd_batch_status(None, None) == _lib.WD_EINVAL
    assert lib.wd_batch_set_spin_limit(None, 0) == _lib.WD_EINVAL
    assert lib.wd_batch_workgroups(None) == _lib.WD_EINVAL
    tile = ctypes.c_int(5)
    assert lib.wd_batch_column_kernel(None, 0, ctypes.byref(tile)) == _lib.WD_EINVAL and tile.value == 5
    assert lib.wd_set_tuned_config(None, -1, 1) == _lib.WD_EINVAL


def test_namelist_validation_matches_supported_path():
    nl = Namelist(dt_s=6.0, acoustic_substeps=6)
    nl.validate()
    for bad in (dict(top_lid=False), dict(hypsometric_opt=3), dict(damp_opt=1), dict(w_damping=2),
                dict(acoustic_substeps=0), dict(run_physics=True), dict(h_sca_adv_order=5),
                dict(const_nu_m2_s=-1.0), dict(const_nu_m2_s=float("inf")),
                dict(const_nu_m2_s=75.0, diff_opt=1, km_opt=4), dict(moist_adv_opt=1),
                dict(moist=True, moist_adv_opt=3)):
        with pytest.raises(ValueError):
            nl.replace(**bad).validate()


def test_metrics_shapes():
    m = Metrics.flat(nx=5, ny=4, nz=6, dx=1.0, dy=1.0, eta=[1.0, 0.8, 0.6, 0.45, 0.3, 0.15, 0.0], p_top=5000.0)
    sh = m.shapes()
    assert sh["u"] == (6, 4, 6) and sh["v"] == (6, 5, 5) and sh["w"] == (7, 4, 5)
    assert m.dnw[0] < 0 and m.rdnw[0] < 0  # WRF-signed metric
    with pytest.raises(ValueError):
        Metrics.flat(nx=5, ny=4, nz=3, dx=1.0, dy=1.0, eta=[1.0, 0.6, 0.3, 0.0], p_top=5000.0)


def test_noalloc_refuses_a_grid_beyond_its_scratch_before_any_device_call():
    """wd_step_noalloc checks its fixed scratch size on the host: a grid that needs more is
    refused with the size before any HIP call (here without a GPU)."""

    from wrf_gpu_amd import ic
    from wrf_gpu_amd.dycore import _abi_structs

    lib = _lib.load()
    case = ic.bubble_3d(nx=96, ny=96, nz=40, noise=0.0)
    keep: list = []
    grid, nl = _abi_structs(case.metrics, case.namelist, keep)
    fake = _lib.WdState(*[ctypes.c_void_p(4096 * (q + 1)) for q in range(11)])  # never dereferenced
    assert lib.wd_step_noalloc(ctypes.byref(grid), ctypes.byref(nl), ctypes.byref(fake), 1, None) == _lib.WD_EINVAL
    msg = _lib.last_error(None)
    assert "scratch" in msg and "wd_create" in msg
    assert lib.wd_step_noalloc(None, None, None, 1, None) == _lib.WD_EINVAL
