"""ctypes loading of the gfx950 libraries (built in-tree by rte-rrtmgp-cpp_amd/build.py).

There is no CPU fallback: if a library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
ROOT = os.path.dirname(PKG_DIR)
INCLUDE_DIR = os.path.join(ROOT, "include")

HIP_LIB = os.path.join(LIB_DIR, "librrtmgp_hip.so")
CPP_LIB = {8: os.path.join(LIB_DIR, "librrtmgp_cpp_f64.so"),
           4: os.path.join(LIB_DIR, "librrtmgp_cpp_f32.so")}

_hip = None
_cpp = {}


def _load(path, mode=C.RTLD_GLOBAL):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run rte-rrtmgp-cpp_amd/build.py "
                           "(the HIP implementation has no fallback)")
    return C.CDLL(path, mode=mode)


def hip():
    """librrtmgp_hip.so: kernels and the thin C ABI (include/rrtmgp_hip.h)."""
    global _hip
    if _hip is None:
        L = _load(HIP_LIB)
        L.rrtmgp_hip_last_error.restype = C.c_char_p
        _hip = L
    return _hip


def cpp(precision=8):
    """librrtmgp_cpp_f64/f32.so: the class API behind the C facade (include/rrtmgp_api.h)."""
    if precision not in _cpp:
        hip()
        L = _load(CPP_LIB[precision], mode=C.RTLD_LOCAL)
        L.rrtmgp_api_last_error.restype = C.c_char_p
        for name in ("rrtmgp_api_kdist_create", "rrtmgp_api_cloud_create", "rrtmgp_api_gas_concs_create",
                     "rrtmgp_api_solver_lw_create", "rrtmgp_api_solver_sw_create"):
            getattr(L, name).restype = C.c_int
        if L.rrtmgp_api_precision() != precision:
            raise RuntimeError("precision mismatch of " + CPP_LIB[precision])
        _cpp[precision] = L
    return _cpp[precision]


class HipError(RuntimeError):
    pass


def check_hip(rc):
    if rc != 0:
        raise HipError(hip().rrtmgp_hip_last_error().decode())


def check_api(rc, precision):
    if rc != 0:
        msg = cpp(precision).rrtmgp_api_last_error().decode()
        if rc == 2:
            raise ValueError(msg)       # std::range_error
        if rc == 1:
            raise TypeError(msg)        # std::invalid_argument
        raise RuntimeError(msg)
