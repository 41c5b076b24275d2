"""The C-ABI libraries load on a machine without a GPU and export every entry point the headers
in include/ declare (no compute calls are made)."""
import os
import re

import pytest

from conftest import ROOT


def _declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    names = set()
    for m in re.finditer(r"\b(" + prefix + r"\w*?)(_##S)?\s*\(", text):
        base, templ = m.group(1), m.group(2)
        if templ:
            names.update({base + "_f32", base + "_f64"})
        else:
            names.add(base)
    return sorted(names)


def test_hip_library_exports_header():
    from rrtmgp_amd import _lib
    L = _lib.hip()
    names = _declared("rrtmgp_hip.h", "rrtmgp_hip_")
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("precision", [4, 8])
def test_api_library_exports_header(precision):
    from rrtmgp_amd import _lib
    L = _lib.cpp(precision)
    names = _declared("rrtmgp_api.h", "rrtmgp_api_")
    assert len(names) > 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.rrtmgp_api_precision() == precision


def test_no_device_reports_error_not_crash():
    import ctypes as C
    from rrtmgp_amd import _lib, device
    if device.device_count() > 0:
        pytest.skip("a device is visible")
    p = C.c_void_p()
    assert _lib.hip().rrtmgp_hip_malloc(C.byref(p), C.c_size_t(16)) != 0
    assert _lib.hip().rrtmgp_hip_last_error()
