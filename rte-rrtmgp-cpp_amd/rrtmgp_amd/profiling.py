"""Per-kernel launch timing of the HIP library (rrtmgp_hip_timing_* in include/rrtmgp_hip.h).

While timing is on, every kernel launch of librrtmgp_hip.so is bracketed by two HIP events on
its own launch stream; collect() waits for them and returns per-kernel totals.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib


@dataclass
class KernelTime:
    name: str
    launches: int
    total_ms: float

    @property
    def avg_ms(self) -> float:
        return self.total_ms / self.launches if self.launches else 0.0


def enable(on: bool = True) -> None:
    """Switch launch timing on (clearing earlier records) or off (keeping them)."""
    _lib.check_hip(_lib.hip().rrtmgp_hip_timing_enable(int(bool(on))))


def collect() -> list[KernelTime]:
    """Totals of the launches recorded since the last enable(True), in first-launch order."""
    L = _lib.hip()
    n = C.c_int(0)
    _lib.check_hip(L.rrtmgp_hip_timing_collect(C.byref(n)))
    out = []
    buf = C.create_string_buffer(128)
    launches, total = C.c_int64(0), C.c_double(0.0)
    for i in range(n.value):
        _lib.check_hip(L.rrtmgp_hip_timing_get(i, buf, len(buf), C.byref(launches), C.byref(total)))
        out.append(KernelTime(buf.value.decode(), int(launches.value), float(total.value)))
    return out
