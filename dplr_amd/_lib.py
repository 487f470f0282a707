"""ctypes binding of libdplr_hip.so (the C-ABI of include/dplr.h and include/dplrx.h).

The shared library holds the GPU implementation; there is no Python or CPU
fallback.  Importing this module raises when the library has not been built.
"""
import ctypes
import os
from ctypes import (POINTER, c_char, c_double, c_float, c_int, c_long, c_size_t, c_uint,
                    c_ulonglong, c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdplr_hip.so")


from .abi import PLR_DMTX, PLR_HMTX, PLR_LRMTX, PLR_UMTX, PD, PM, DPlrMtx, declare  # noqa: F401


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(make -C dplr_amd/csrc).  There is no CPU fallback.")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, the same soname as /opt/rocm's).  Loading torch first
    # makes the dynamic loader resolve this library's libamdhip64.so.7 to
    # torch's copy, so device pointers from torch tensors are valid here and
    # no second HSA runtime competes for the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    declare(lib)
    P = c_void_p
    x = {
        "dplrx_plan_create": (c_int, [c_int, PM, POINTER(P)]),
        "dplrx_plan_destroy": (None, [P]),
        "dplrx_plan_info": (c_int, [P, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "dplrx_plan_bytes": (c_size_t, [P]),
        "dplrx_gepmv": (c_int, [P, c_char, c_double, P, c_double, P, P]),
        "dplrx_gepmv_noalloc": (c_int, [P, c_char, c_double, P, c_double, P, P]),
        "dplrx_gepdmm": (c_int, [P, c_char, c_int, c_double, P, c_int, c_double, P, c_int, P]),
        "dplrx_last_error": (ctypes.c_char_p, []),
        "dplrx_factor": (c_int, [P, POINTER(c_int), P]),
        "dplrx_solve": (c_int, [P, c_char, c_int, P, c_int, P, c_int, P]),
        "dplrx_factor_mixed": (c_int, [P, c_int, POINTER(c_int), P]),
        "dplrx_solve_refined": (c_int, [P, c_int, P, c_int, P, c_int, c_int, PD, P]),
        "dplrx_solve_refined_tol": (c_int, [P, c_int, P, c_int, P, c_int, c_double, c_int,
                                            POINTER(c_int), PD, P]),
        "dplrx_apply_solve_refined_tol": (c_int, [P, c_int, P, c_int, P, c_int, P, c_int, P,
                                                  c_int, c_double, c_int, POINTER(c_int), PD,
                                                  P]),
        "dplrx_grad_outer": (c_int, [P, c_int, c_double, P, c_int, P, c_int, P, P, P, P, P]),
        "dplrx_batch_create": (c_int, [c_int, POINTER(P), POINTER(P)]),
        "dplrx_batch_destroy": (None, [P]),
        "dplrx_batch_gepdmm": (c_int, [P, c_int, POINTER(P), POINTER(c_int), POINTER(P),
                                       POINTER(c_int), P]),
        "dplrx_batch_solve": (c_int, [P, c_int, POINTER(P), POINTER(c_int), POINTER(P),
                                      POINTER(c_int), P]),
        "dplrx_batch_set_tile": (c_int, [P, c_int]),
        "dplrx_batch_tune": (c_int, [P, c_int, POINTER(P), POINTER(c_int), POINTER(P),
                                     POINTER(c_int), c_int, POINTER(c_int), PD, P]),
        "dplrx_factor_single": (c_int, [P, c_int, POINTER(c_int), P]),
        "dplrx_gepdmm_single": (c_int, [P, c_int, c_char, c_int, c_float, P, c_int, c_float, P,
                                        c_int, P]),
        "dplrx_solve_single": (c_int, [P, c_int, c_char, c_int, P, c_int, P, c_int, P]),
        "dplrx_grad_outer_f32": (c_int, [P, c_int, c_float, P, c_int, P, c_int, P, P, P, P, P]),
        "dplrx_batch_op": (c_int, [P, c_int, c_int, c_char, c_int, POINTER(P), POINTER(c_int),
                                   POINTER(P), POINTER(c_int), c_double, c_double, P]),
        "dplrx_batch_solve_refined": (c_int, [P, c_int, POINTER(P), POINTER(c_int), POINTER(P),
                                              POINTER(c_int), c_int, P]),
        "dplrx_batch_tune_prec": (c_int, [P, c_int, c_int, POINTER(P), POINTER(c_int), POINTER(P),
                                          POINTER(c_int), c_int, POINTER(c_int), PD, P]),
        "dplrx_gen_random": (c_int, [c_int, c_int, c_int, c_double, c_ulonglong, PM]),
        "dplrx_save": (c_int, [ctypes.c_char_p, c_int, PM]),
        "dplrx_load": (c_int, [ctypes.c_char_p, PM, POINTER(c_int)]),
        "dplrx_gen_normal": (c_int, [c_size_t, c_ulonglong, c_uint, PD]),
    }
    for name, (res, args) in x.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib


lib = _load()


def check(rc, what="dplrx call"):
    if rc != 0:
        msg = lib.dplrx_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
