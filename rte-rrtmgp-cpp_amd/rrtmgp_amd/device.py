"""Device arrays owned through the C ABI (rrtmgp_hip_malloc / memcpy), Fortran layout."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A device buffer holding an array in Fortran order (first index fastest)."""

    def __init__(self, shape, dtype=np.float64, zero=True):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        _lib.check_hip(_lib.hip().rrtmgp_hip_malloc(C.byref(p), C.c_size_t(max(self.nbytes, 1))))
        self._p = p
        if zero:
            _lib.check_hip(_lib.hip().rrtmgp_hip_memset(p, 0, C.c_size_t(max(self.nbytes, 1)), None))

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.asfortranarray(np.asarray(a, dtype=dtype))
        d = cls(a.shape, a.dtype, zero=False)
        if a.size:
            _lib.check_hip(_lib.hip().rrtmgp_hip_memcpy_h2d(d._p, a.ctypes.data_as(C.c_void_p),
                                                            C.c_size_t(d.nbytes), None))
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.size:
            _lib.check_hip(_lib.hip().rrtmgp_hip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self._p,
                                                            C.c_size_t(self.nbytes), None))
        return out

    @property
    def ptr(self):
        return self._p

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            _lib.hip().rrtmgp_hip_free(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _lib.check_hip(_lib.hip().rrtmgp_hip_device_synchronize())


def device_count():
    n = C.c_int(0)
    rc = _lib.hip().rrtmgp_hip_device_count(C.byref(n))
    return n.value if rc == 0 else 0


class Event:
    def __init__(self):
        self._e = C.c_void_p()
        _lib.check_hip(_lib.hip().rrtmgp_hip_event_create(C.byref(self._e)))

    def record(self, stream=None):
        _lib.check_hip(_lib.hip().rrtmgp_hip_event_record(self._e, stream))

    def elapsed_ms(self, end):
        ms = C.c_float(0)
        _lib.check_hip(_lib.hip().rrtmgp_hip_event_elapsed_ms(self._e, end._e, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            _lib.hip().rrtmgp_hip_event_destroy(self._e)
        except Exception:
            pass
