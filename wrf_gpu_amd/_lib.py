"""ctypes binding of libwrfdyn.so (the C ABI in include/wrfdyn.h).

There is no fallback: if the HIP library is missing this raises, so a GPU
run can never silently take another path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwrfdyn.so")

_dp = ctypes.POINTER(ctypes.c_double)

WD_OK = 0
WD_EINVAL = -1
WD_ENOMEM = -2
WD_EDEVICE = -3

WD_PREC_FP64 = 0
WD_PREC_MIXED = 1

EXPORTED_SYMBOLS = ("wd_create", "wd_step", "wd_workspace_bytes", "wd_last_error", "wd_destroy",
                    "wd_profile_enable", "wd_profile_read", "wd_kernel_count", "wd_kernel_name",
                    "wd_capture_graph", "wd_graph_step", "wd_autotune", "wd_tuned_config", "wd_set_tuned_config",
                    "wd_set_precision", "wd_get_precision",
                    "wd_batch_create", "wd_step_batched", "wd_batch_branch", "wd_batch_branches",
                    "wd_batch_last_error", "wd_batch_destroy", "wd_step_adjoint", "wd_adjoint_tape_bytes",
                    "wd_step_noalloc", "wd_batch_create_persistent", "wd_batch_status", "wd_batch_set_spin_limit",
                    "wd_batch_workgroups", "wd_batch_column_kernel")


class WdGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("c1h", _dp), ("c2h", _dp), ("c1f", _dp), ("c2f", _dp),
        ("dnw", _dp), ("rdnw", _dp), ("dn", _dp), ("rdn", _dp), ("fnm", _dp), ("fnp", _dp),
        ("cf1", ctypes.c_double), ("cf2", ctypes.c_double), ("cf3", ctypes.c_double),
        ("msftx", _dp), ("msfty", _dp), ("msfux", _dp), ("msfuy", _dp), ("msfvx", _dp), ("msfvy", _dp),
        ("c3h", _dp), ("c4h", _dp), ("c3f", _dp), ("c4f", _dp), ("p_top", ctypes.c_double),
        ("f", _dp), ("e", _dp), ("sina", _dp), ("cosa", _dp),
    ]


class WdNamelist(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("acoustic_substeps", ctypes.c_int), ("epssm", ctypes.c_double),
        ("top_lid", ctypes.c_int), ("w_damping", ctypes.c_int), ("damp_opt", ctypes.c_int),
        ("dampcoef", ctypes.c_double), ("zdamp", ctypes.c_double),
        ("diff_6th_opt", ctypes.c_int), ("diff_6th_factor", ctypes.c_double),
        ("diff_opt", ctypes.c_int), ("km_opt", ctypes.c_int), ("c_s", ctypes.c_double),
        ("scalar_adv_opt", ctypes.c_int), ("hypsometric_opt", ctypes.c_int),
        ("const_nu_m2_s", ctypes.c_double), ("moist", ctypes.c_int), ("moist_adv_opt", ctypes.c_int),
    ]


class WdState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "u", "v", "w", "theta", "p_total", "p_pert", "ph_total", "ph_pert", "mu_total", "mu_pert", "mudf",
        "qv", "qc", "qr", "qi", "qs", "qg")]


_lib = None


class WrfdynError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"wrfdyn error {code}: {message}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load libwrfdyn.so; raises if it has not been built."""

    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as e; e.build()')")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7) as the system ROCm: whichever loads first serves
        # the whole process.  Import torch before the library so that both
        # use torch's; loaded the other way round, torch finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        lib.wd_create.restype = ctypes.c_int
        lib.wd_create.argtypes = [ctypes.POINTER(WdGrid), ctypes.POINTER(WdNamelist),
                                  ctypes.POINTER(ctypes.c_void_p)]
        lib.wd_step.restype = ctypes.c_int
        lib.wd_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(WdState), ctypes.c_int, ctypes.c_void_p]
        lib.wd_workspace_bytes.restype = ctypes.c_size_t
        lib.wd_workspace_bytes.argtypes = [ctypes.c_void_p]
        lib.wd_last_error.restype = ctypes.c_char_p
        lib.wd_last_error.argtypes = [ctypes.c_void_p]
        lib.wd_destroy.restype = None
        lib.wd_destroy.argtypes = [ctypes.c_void_p]
        lib.wd_profile_enable.restype = ctypes.c_int
        lib.wd_profile_enable.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        lib.wd_profile_read.restype = ctypes.c_int
        lib.wd_profile_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        lib.wd_kernel_count.restype = ctypes.c_int
        lib.wd_kernel_count.argtypes = []
        lib.wd_kernel_name.restype = ctypes.c_char_p
        lib.wd_kernel_name.argtypes = [ctypes.c_int]
        lib.wd_capture_graph.restype = ctypes.c_int
        lib.wd_capture_graph.argtypes = [ctypes.c_void_p, ctypes.POINTER(WdState)]
        lib.wd_graph_step.restype = ctypes.c_int
        lib.wd_graph_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.wd_step_adjoint.restype = ctypes.c_int
        lib.wd_step_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(WdState), ctypes.POINTER(WdState),
                                        ctypes.c_void_p]
        lib.wd_adjoint_tape_bytes.restype = ctypes.c_size_t
        lib.wd_adjoint_tape_bytes.argtypes = [ctypes.c_void_p]
        lib.wd_autotune.restype = ctypes.c_int
        lib.wd_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(WdState), ctypes.c_int, ctypes.c_void_p]
        lib.wd_set_precision.restype = ctypes.c_int
        lib.wd_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.wd_get_precision.restype = ctypes.c_int
        lib.wd_get_precision.argtypes = [ctypes.c_void_p]
        lib.wd_tuned_config.restype = ctypes.c_int
        lib.wd_tuned_config.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.wd_set_tuned_config.restype = ctypes.c_int
        lib.wd_set_tuned_config.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.wd_batch_create.restype = ctypes.c_int
        lib.wd_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(WdState), ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        lib.wd_step_batched.restype = ctypes.c_int
        lib.wd_step_batched.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        lib.wd_batch_branch.restype = ctypes.c_int
        lib.wd_batch_branch.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.wd_batch_branches.restype = ctypes.c_int
        lib.wd_batch_branches.argtypes = [ctypes.c_void_p]
        lib.wd_batch_last_error.restype = ctypes.c_char_p
        lib.wd_batch_last_error.argtypes = [ctypes.c_void_p]
        lib.wd_step_noalloc.restype = ctypes.c_int
        lib.wd_step_noalloc.argtypes = [ctypes.POINTER(WdGrid), ctypes.POINTER(WdNamelist), ctypes.POINTER(WdState),
                                        ctypes.c_int, ctypes.c_void_p]
        lib.wd_batch_destroy.restype = None
        lib.wd_batch_destroy.argtypes = [ctypes.c_void_p]
        lib.wd_batch_create_persistent.restype = ctypes.c_int
        lib.wd_batch_create_persistent.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(WdState),
                                                   ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        lib.wd_batch_status.restype = ctypes.c_int
        lib.wd_batch_status.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.wd_batch_set_spin_limit.restype = ctypes.c_int
        lib.wd_batch_set_spin_limit.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        lib.wd_batch_workgroups.restype = ctypes.c_int
        lib.wd_batch_workgroups.argtypes = [ctypes.c_void_p]
        lib.wd_batch_column_kernel.restype = ctypes.c_int
        lib.wd_batch_column_kernel.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        _lib = lib
    return _lib


def kernel_names() -> list:
    lib = load()
    return [lib.wd_kernel_name(i).decode() for i in range(lib.wd_kernel_count())]


def last_error(handle=None) -> str:
    msg = load().wd_last_error(handle)
    return msg.decode() if msg else ""


def check(code: int, handle=None) -> None:
    if code != WD_OK:
        raise WrfdynError(code, last_error(handle))
