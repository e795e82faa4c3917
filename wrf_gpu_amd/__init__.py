"""MI355X-native split-explicit ARW dycore large step (wrf_gpu hot path).

Host mirror of the reference's dycore step API over hand-written HIP kernels
(libwrfdyn.so, C ABI in include/wrfdyn.h).
"""

from .grid import Metrics, Namelist, STATE_FIELDS  # noqa: F401

__all__ = ["Metrics", "Namelist", "STATE_FIELDS"]
