"""MI355X-native PLR (HODLR) apply/solve: the dplr C API on hand-written gfx950 kernels.

`dplr_amd/libdplr_hip.so` exports the reference's C entry points (include/dplr.h)
and a device-resident extension (include/dplrx.h); this package is a thin
ctypes mirror of both for Python callers and tests.
"""
from ._lib import LIB_PATH, lib  # noqa: F401
from .tree import PlrTree, geinv, gepdmm, gepmv, geprk  # noqa: F401
from .plan import Batch, Plan  # noqa: F401
