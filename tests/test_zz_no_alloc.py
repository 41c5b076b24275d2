"""The shortwave solver entry point that allocates nothing (rrtmgp_hip_sw_solver_2stream_bnd_noalloc_*)
against the allocating one (rrtmgp_hip_sw_solver_2stream_bnd_*), through the C ABI.

Same kernels, same launch shapes: band fluxes, broadband fluxes and heating rates must be equal
bit for bit.  Two calls in a row, the second on fewer columns and layers, so that the second runs
on scratch still holding the first call's values; then a call beyond the scratch size must be
refused with an error, not run.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import AVAILABLE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from rrtmgp_amd import api, device
    if device.device_count() < 1:
        pytest.fail("no HIP device visible")
    return api


def _inputs(G, O, syn, okd_sw, lut_sw, ncol, nlay, seed, dt):
    atm = syn.make_atmosphere(seed, ncol, nlay, cloudy=True)
    go = O.gas_optics(okd_sw, atm)
    ct, cs, cg = O.delta_scale(*O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    tau, ssa, g = O.inc_bybnd(go["tau"], ct, okd_sw.band_lims, go["ssa"], go["g"], cs, cg)
    dev = lambda a: G.DeviceArray.from_host(np.asfortranarray(a), dt)     # noqa: E731
    return dict(tau=dev(tau), ssa=dev(ssa), g=dev(g), mu0=dev(atm.mu0), toa=dev(go["toa_src"]),
                alb_dir=dev(atm.sfc_alb), alb_dif=dev(0.5 * atm.sfc_alb), plev=dev(atm.p_lev),
                top_at_1=int(atm.top_at_1))


def _solve(G, entry, d, ncol, nlay, ngpt, bl, dt):
    nbnd = bl.shape[1]
    nlev = nlay + 1
    out = {k: G.DeviceArray((ncol, nlev, nbnd), dt) for k in ("bu", "bd", "bdir")}
    out.update({k: G.DeviceArray((ncol, nlev), dt) for k in ("u", "dn", "dir")})
    out["hr"] = G.DeviceArray((ncol, nlay), dt)
    i = C.c_int64
    rc = entry(i(ncol), i(nlay), i(ngpt), C.c_int(d["top_at_1"]), d["tau"].ptr, d["ssa"].ptr, d["g"].ptr,
               d["mu0"].ptr, d["alb_dir"].ptr, d["alb_dif"].ptr, d["toa"].ptr, C.c_int(0), None, i(nbnd),
               bl.ctypes.data_as(C.POINTER(C.c_int)), out["bu"].ptr, out["bd"].ptr, out["bdir"].ptr,
               out["u"].ptr, out["dn"].ptr, out["dir"].ptr, d["plev"].ptr, out["hr"].ptr, None)
    return rc, out


@pytest.mark.parametrize("prec", [8, 4])
def test_sw_solver_without_allocation_matches(G, O, syn, okd_sw, lut_sw, prec):
    from rrtmgp_amd import _lib
    L = _lib.hip()
    S = "f64" if prec == 8 else "f32"
    dt = np.float64 if prec == 8 else np.float32
    plain = getattr(L, "rrtmgp_hip_sw_solver_2stream_bnd_" + S)
    noalloc = getattr(L, "rrtmgp_hip_sw_solver_2stream_bnd_noalloc_" + S)
    bl = np.asfortranarray(np.asarray(okd_sw.band_lims, dtype=np.int32))
    ngpt = int(bl.max())
    for ncol, nlay, seed in ((150, 72, 81), (37, 40, 82)):
        d = _inputs(G, O, syn, okd_sw, lut_sw, ncol, nlay, seed, dt)
        rc, want = _solve(G, plain, d, ncol, nlay, ngpt, bl, dt)
        _lib.check_hip(rc)
        rc, got = _solve(G, noalloc, d, ncol, nlay, ngpt, bl, dt)
        _lib.check_hip(rc)
        for k in want:
            a, b = got[k].to_host(), want[k].to_host()
            assert np.array_equal(a, b), (ncol, nlay, k, float(np.abs(a - b).max()))


def test_sw_solver_without_allocation_refuses_oversize(G):
    """A call whose scratch exceeds the fixed array returns an error before any launch."""
    from rrtmgp_amd import _lib
    L = _lib.hip()
    ncol, nlay, ngpt = 100000, 128, 224
    bl = np.asfortranarray(np.array([[1 + 16 * b for b in range(14)], [16 * (b + 1) for b in range(14)]],
                                    dtype=np.int32))
    one = G.DeviceArray((1,), np.float32)
    i = C.c_int64
    p = one.ptr
    rc = L.rrtmgp_hip_sw_solver_2stream_bnd_noalloc_f32(
        i(ncol), i(nlay), i(ngpt), C.c_int(0), p, p, p, p, p, p, p, C.c_int(0), None, i(14),
        bl.ctypes.data_as(C.POINTER(C.c_int)), None, None, None, p, p, p, None, None, None)
    assert rc != 0
    assert "bytes of scratch" in L.rrtmgp_hip_last_error().decode()
