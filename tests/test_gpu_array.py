"""Array_gpu element operations behind the C ABI: rrtmgp_hip_fill and rrtmgp_hip_subset.

The counterparts of the reference's fill_kernel / subset_kernel (include/Array.h:296-335): checked
against numpy slicing, including the spread of a length-1 dimension over its range
(Array.h:230-233), Array offsets, 1-, 4- and 8-byte elements, and the rejection of a range that
leaves the source array (which the reference would read out of bounds).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _i64(v):
    a = np.asarray(v, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


def test_fill():
    from rrtmgp_amd import _lib
    from rrtmgp_amd.device import DeviceArray, synchronize
    L = _lib.hip()
    for dt, val in ((np.float64, -2.5), (np.float32, 3.25), (np.int32, -7), (np.int8, 5)):
        d = DeviceArray((1000, 3), dt, zero=False)
        v = np.array(val, dtype=dt)
        print(f"fill {np.dtype(dt).name}", flush=True)
        _lib.check_hip(L.rrtmgp_hip_fill(d.ptr, C.c_int64(3000), C.c_int(v.itemsize),
                                         v.ctypes.data_as(C.c_void_p), None))
        synchronize()
        assert np.all(d.to_host() == v)


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int8])
def test_subset(dt):
    from rrtmgp_amd import _lib
    from rrtmgp_amd.device import DeviceArray, synchronize
    L = _lib.hip()
    rng = np.random.default_rng(3)
    src = np.asfortranarray((rng.random((9, 1, 7)) * 100).astype(dt))   # dim 2 has length 1
    d_src = DeviceArray.from_host(src)
    # 1-based inclusive ranges {3..8}, {1..4} (spread), {2..6}
    starts = [3, 1, 2]
    sub_dims = [6, 4, 5]
    d_dst = DeviceArray(tuple(sub_dims), dt, zero=False)
    sd, p_sd = _i64(src.shape)
    st, p_st = _i64(starts)
    dd, p_dd = _i64(sub_dims)
    print(f"subset {np.dtype(dt).name}", flush=True)
    _lib.check_hip(L.rrtmgp_hip_subset(C.c_int(src.itemsize), C.c_int(3), d_src.ptr, p_sd, None, p_st,
                                       d_dst.ptr, p_dd, None))
    synchronize()
    want = np.repeat(src[2:8, 0:1, 1:6], 4, axis=1)
    assert np.array_equal(d_dst.to_host(), want)

    # Offsets: with offsets (-1, 0, 2) index i of dim k is element i - offset_k.
    of, p_of = _i64([-1, 0, 2])
    st2, p_st2 = _i64([1, 1, 4])        # elements 2.., 1, 2..
    d2 = DeviceArray((3, 1, 3), dt, zero=False)
    dd2, p_dd2 = _i64([3, 1, 3])
    _lib.check_hip(L.rrtmgp_hip_subset(C.c_int(src.itemsize), C.c_int(3), d_src.ptr, p_sd, p_of, p_st2,
                                       d2.ptr, p_dd2, None))
    synchronize()
    assert np.array_equal(d2.to_host(), src[1:4, 0:1, 1:4])

    # A range past the end of the source is an argument error, not a launch.
    bad, p_bad = _i64([5, 1, 1])
    dd3, p_dd3 = _i64([6, 1, 1])
    rc = L.rrtmgp_hip_subset(C.c_int(src.itemsize), C.c_int(3), d_src.ptr, p_sd, None, p_bad, d_dst.ptr,
                             p_dd3, None)
    assert rc == 1
