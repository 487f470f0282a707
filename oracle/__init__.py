"""TEST INFRASTRUCTURE: loaders for the CPU oracle and the reference build.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker.  The product (dplr_amd/) never imports it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libdplr_ref.so")


def _declare(lib, prefix):
    from dplr_amd.abi import declare
    return declare(lib, prefix)


def load_oracle():
    """dict name -> function for the clean-room restatement (orc_ symbols)."""
    if not os.path.exists(ORACLE_SO):
        raise ImportError(f"{ORACLE_SO} missing: run make -C oracle")
    lib = ctypes.CDLL(ORACLE_SO, mode=ctypes.RTLD_LOCAL)
    api = _declare(lib, "orc_")
    api["_lib"] = lib
    return api


def load_ref():
    """dict name -> function for the reference sources built into oracle/_ref,
    or None when that build is absent."""
    if not os.path.exists(REF_SO):
        return None
    lib = ctypes.CDLL(REF_SO, mode=ctypes.RTLD_LOCAL)
    api = _declare(lib, "")
    api["_lib"] = lib
    return api
