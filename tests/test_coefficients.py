"""Coefficient and case schemas (the reference's NetCDF names and dimension orders) and bundles.

A k-distribution or cloud table written as a bundle and read back must give identical arrays,
and its dimension lengths must be the ones the reference's reader asks for
(src_test/Radiation_solver.cpp:70-327).  The C++ reader of the same bundles is exercised by
tests/test_gpu_drop_in.py.
"""
import numpy as np

from rrtmgp_amd import coefficients as co


def _same(a, b):
    if isinstance(a, list):
        return a == b
    if a is None or b is None:
        return a is b
    if isinstance(a, float):
        return a == b
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_kdist_bundle_roundtrip(tmp_path, kd_lw, kd_sw):
    for kd, name in ((kd_lw, "lw"), (kd_sw, "sw")):
        co.write_bundle(str(tmp_path / name), co.kdist_schema(kd))
        s = co.read_bundle(str(tmp_path / name))
        back = co.kdist_from_schema(s)
        for f in kd.__dataclass_fields__:
            assert _same(getattr(kd, f), getattr(back, f)), (name, f)
        # The dimensions the reference reader asks for.
        assert s.dims["gpt"] == kd.ngpt and s.dims["bnd"] == kd.nbnd
        assert s.dims["pressure"] == kd.press_ref.size and s.dims["temperature"] == kd.temp_ref.size
        assert s.vars["kmajor"][0] == ("temperature", "pressure_interp", "mixing_fraction", "gpt")
        # C order over the NetCDF dims is the Fortran order of the reversed dims.
        k = s.get("kmajor")
        assert k[3, 10, 2, 17] == kd.kmajor[17, 2, 10, 3]


def test_cloud_bundle_roundtrip(tmp_path, lut_sw):
    co.write_bundle(str(tmp_path / "c"), co.cloud_schema(lut_sw))
    back = co.cloud_from_schema(co.read_bundle(str(tmp_path / "c")))
    for f in lut_sw.__dataclass_fields__:
        assert _same(getattr(lut_sw, f), getattr(back, f)), f


def test_case_roundtrip(tmp_path, syn):
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 7, 12, cloudy=True)
    co.write_bundle(str(tmp_path / "case"), co.case_schema(atm, tsi_scaling=1.0))
    s = co.read_bundle(str(tmp_path / "case"))
    assert s.vars["p_lay"][0] == ("lay", "y", "x") and s.vars["emis_sfc"][0] == ("y", "x", "band_lw")
    back = co.atmosphere_from_case(s)
    for f in ("p_lay", "p_lev", "t_lay", "t_lev", "t_sfc", "emis_sfc", "sfc_alb", "mu0", "lwp", "iwp",
              "rel", "rei"):
        assert np.array_equal(getattr(atm, f), getattr(back, f)), f
    for g, v in atm.vmr.items():
        assert np.array_equal(np.broadcast_to(v, back.vmr[g].shape), back.vmr[g]), g
    assert back.top_at_1 == atm.top_at_1
