"""The C ABI of include/dplr.h described with ctypes (no library is loaded here).

The node layout is the reference's (include/dplr_internal.h:32-87), so the
same description serves the product library, the oracle and the reference
build used by the tests.
"""
import ctypes
from ctypes import POINTER, c_char, c_double, c_int, c_long

class DPlrMtx(ctypes.Structure):
    """The 40-byte node of include/dplr.h (reference include/dplr_internal.h:75-87)."""


class _HNode(ctypes.Structure):
    _fields_ = [("nw", POINTER(DPlrMtx)), ("ne", POINTER(DPlrMtx)),
                ("se", POINTER(DPlrMtx)), ("sw", POINTER(DPlrMtx))]


class _DNode(ctypes.Structure):
    _fields_ = [("d", POINTER(c_double))]


class _LRNode(ctypes.Structure):
    _fields_ = [("rank", c_int), ("capacity", c_int), ("u", POINTER(c_double)),
                ("v", POINTER(c_double)), ("sig", POINTER(c_double))]


class _Body(ctypes.Union):
    _fields_ = [("dh", _HNode), ("dd", _DNode), ("dlr", _LRNode)]


DPlrMtx._anonymous_ = ("body",)
DPlrMtx._fields_ = [("type", c_int), ("size", c_int), ("body", _Body)]
assert ctypes.sizeof(DPlrMtx) == 40

PLR_UMTX, PLR_HMTX, PLR_DMTX, PLR_LRMTX = 0, 1, 2, 3

PM = POINTER(DPlrMtx)
PD = POINTER(c_double)


def declare(lib, prefix=""):
    """Set argument and result types of the dplr.h API on a library object.

    `prefix` lets the same table describe the oracle (orc_ names)."""
    sig = {
        "dplrini": (None, [PM]),
        "dplrclr": (None, [PM]),
        "dplrinid": (None, [c_int, PM]),
        "dplrinilr": (None, [c_int, PM, c_int]),
        "dplrinih": (None, [c_int, PM]),
        "dplrresizelr": (None, [PM, c_int]),
        "dplrgep2d": (None, [c_int, PM, PD, c_int]),
        "dplrgepmv": (None, [c_char, c_int, PD, PM, PD, c_int, PD, PD, c_int]),
        "dplrgepdmm": (None, [c_char, c_char, c_int, c_int, PD, PM, PD, c_int, PD, PD, c_int]),
        "dplrgeinv": (None, [c_int, PM, PM, c_double, POINTER(c_int)]),
        "dplrgeprk": (None, [c_int, c_int, PD, PD, c_int, PD, c_int, PD, c_int, PM, c_double]),
        "dplrrnk": (None, [c_int, PM, POINTER(c_int)]),
        "dplrnz": (c_long, [c_int, PM]),
        "dplrd2gep": (None, [c_int, PD, c_int, PM, c_double, c_int, POINTER(c_int)]),
        "dplrgeppmm": (None, [c_char, c_char, c_int, PD, PM, PM, PD, PM, c_double]),
        "dplrempty": (None, [PM, PM]),
        "dplrcpy": (None, [c_int, PM, PM]),
    }
    found = {}
    for name, (res, args) in sig.items():
        full = prefix + name
        if not hasattr(lib, full):
            continue
        f = getattr(lib, full)
        f.restype = res
        f.argtypes = args
        found[name] = f
    return found


