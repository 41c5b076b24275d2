"""The stand-alone driver (rrtmgp_amd.driver, test_rte_rrtmgp.cpp:168-611) end to end: a case and
coefficient files written as NetCDF in a scratch directory, the driver run with the reference's
options, and rte_rrtmgp_output.nc read back and checked against the FP64 oracle.

Bars: FP64 fluxes within 1e-10 relative (floor 1e-6 W m-2), optical properties within 1e-12
relative (1e-6 floor of each field's maximum); FP32 fluxes within 0.05 W m-2.  Dimensions,
variable names, dimension orders and types follow the reference's output file.
"""
import os

import numpy as np
import pytest

from conftest import AVAILABLE, ROOT, rel_err

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def case_dir(tmp_path_factory, syn, kd_lw, kd_sw, lut_lw, lut_sw):
    from rrtmgp_amd import coefficients, netcdf
    d = tmp_path_factory.mktemp("driver_case")
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 37, 60, cloudy=True)
    s = coefficients.case_schema(atm, tsi_scaling=0.98)
    rng = np.random.default_rng(9)
    s.put("rh", ("lay", "y", "x"), rng.uniform(0, 1, (atm.nlay, 1, atm.ncol)))
    for i in range(1, 12):
        if i % 2:
            s.put(f"aermr{i:02d}", ("lay",), rng.uniform(1e-10, 1e-7, atm.nlay))
        else:
            s.put(f"aermr{i:02d}", ("lay", "y", "x"), rng.uniform(1e-10, 1e-7, (atm.nlay, 1, atm.ncol)))
    netcdf.write_schema(str(d / "rte_rrtmgp_input.nc"), s)
    netcdf.write_schema(str(d / "coefficients_lw.nc"), coefficients.kdist_schema(kd_lw))
    netcdf.write_schema(str(d / "coefficients_sw.nc"), coefficients.kdist_schema(kd_sw))
    netcdf.write_schema(str(d / "cloud_coefficients_lw.nc"), coefficients.cloud_schema(lut_lw))
    netcdf.write_schema(str(d / "cloud_coefficients_sw.nc"), coefficients.cloud_schema(lut_sw))
    with open(os.path.join(GOLD, "aerosol_optics.nc"), "rb") as src:
        (d / "aerosol_optics.nc").write_bytes(src.read())
    atm.extra["rh"] = np.asfortranarray(s.get("rh").reshape(atm.nlay, atm.ncol).T)
    atm.extra["aermr"] = {f"aermr{i:02d}": np.asfortranarray(
        np.broadcast_to(s.get(f"aermr{i:02d}").reshape(atm.nlay, -1).T, (atm.ncol, atm.nlay)))
        for i in range(1, 12)}
    return d, atm


def _run(d, *opts):
    from rrtmgp_amd import driver, netcdf
    print("driver", " ".join(opts), flush=True)
    assert driver.solve_radiation(list(opts), workdir=str(d)) == 0
    return netcdf.Dataset(str(d / "rte_rrtmgp_output.nc"))


def _f(ds, name, ncol):
    """An output variable ([band|gpt,] lev|lay, y, x) as the Fortran (ncol, nz[, n]) array."""
    a = ds.get_variable(name).astype(np.float64)
    return np.asfortranarray(a.reshape(a.shape[:-2] + (ncol,)).T)


def test_driver_clear_sky_fp64(case_dir, O, okd_lw, okd_sw):
    d, atm = case_dir
    ds = _run(d)
    assert ds.dimensions["x"] == atm.ncol and ds.dimensions["y"] == 1
    assert ds.variables["lw_flux_up"].dimensions == ("lev", "y", "x")
    assert ds.get_variable("lw_flux_up").dtype == np.float64
    assert "lw_tau" not in ds.variables and "lw_bnd_flux_up" not in ds.variables
    rlw = O.solve_lw(okd_lw, atm)
    rsw = O.solve_sw(okd_sw, atm, tsi_scaling=np.full(atm.ncol, 0.98))
    for k in ("flux_up", "flux_dn", "flux_net"):
        assert rel_err(_f(ds, "lw_" + k, atm.ncol), rlw[k], floor=1e-6) < 1e-10, k
    for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net"):
        assert rel_err(_f(ds, "sw_" + k, atm.ncol), rsw[k], floor=1e-6) < 1e-10, k
    assert np.array_equal(ds.get_variable("lw_band_lims_wvn"), np.asarray(okd_lw.src.band_lims_wavenum).T)
    assert np.array_equal(ds.get_variable("sw_band_lims_wvn"), np.asarray(okd_sw.src.band_lims_wavenum).T)


def test_driver_cloudy_optical_and_band_fluxes(case_dir, O, okd_lw, okd_sw, lut_lw, lut_sw):
    d, atm = case_dir
    ds = _run(d, "--cloud-optics", "--output-optical", "--output-bnd-fluxes")
    n = atm.ncol
    rlw = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw), bnd=True)
    tsi = np.full(n, 0.98)
    rsw = O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), bnd=True, tsi_scaling=tsi)
    for k in ("flux_up", "flux_dn", "flux_net", "bnd_flux_up", "bnd_flux_dn", "bnd_flux_net"):
        assert rel_err(_f(ds, "lw_" + k, n), rlw[k], floor=1e-6) < 1e-10, k
    for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net", "bnd_flux_up", "bnd_flux_dn",
              "bnd_flux_dn_dir", "bnd_flux_net"):
        assert rel_err(_f(ds, "sw_" + k, n), rsw[k], floor=1e-6) < 1e-10, k
    assert ds.variables["lw_bnd_flux_up"].dimensions == ("band_lw", "lev", "y", "x")
    # Optical properties after the cloud additions, as the reference stores them.
    go = O.gas_optics(okd_lw, atm)
    ct = O.Cloud(lut_lw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, False)
    tau = O.inc_bybnd(go["tau"], ct, okd_lw.band_lims)
    assert ds.variables["lw_tau"].dimensions == ("gpt_lw", "lay", "y", "x")
    for name, ref in (("lw_tau", tau), ("lay_source", go["lay_src"]), ("lev_source_inc", go["lev_src_inc"]),
                      ("lev_source_dec", go["lev_src_dec"]), ("sfc_source", go["sfc_src"])):
        assert rel_err(_f(ds, name, n), ref, floor=1e-6 * np.abs(ref).max()) < 1e-12, name
    gs = O.gas_optics(okd_sw, atm)
    t, s, g = O.delta_scale(*O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    tau, ssa, gg = O.inc_bybnd(gs["tau"], t, okd_sw.band_lims, gs["ssa"], gs["g"], s, g)
    for name, ref in (("sw_tau", tau), ("ssa", ssa), ("g", gg), ("toa_source", gs["toa_src"] * tsi[:, None])):
        assert rel_err(_f(ds, name, n), ref, floor=1e-6 * np.abs(ref).max()) < 1e-12, name
    assert np.array_equal(ds.get_variable("sw_band_lims_gpt"), np.asarray(okd_sw.band_lims).T)
    assert ds.get_variable("sw_band_lims_gpt").dtype == np.int32


def test_driver_aerosols_fp64(case_dir, O, okd_sw, lut_sw):
    from rrtmgp_amd import coefficients, netcdf
    d, atm = case_dir
    ds = _run(d, "--no-longwave", "--cloud-optics", "--aerosol-optics", "--output-optical")
    assert "lw_flux_up" not in ds.variables
    n = atm.ncol
    lut = coefficients.aerosol_lut(netcdf.Dataset(str(d / "aerosol_optics.nc")))
    gs = O.gas_optics(okd_sw, atm)
    t, s, g = O.delta_scale(*O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    bl = okd_sw.band_lims
    tau, ssa, gg = O.inc_bybnd(gs["tau"], t, bl, gs["ssa"], gs["g"], s, g)
    at, as_, ag = O.aerosol_optics(atm.extra["aermr"], atm.extra["rh"], atm.p_lev, lut)
    tau, ssa, gg = O.inc_bybnd(tau, at, bl, ssa, gg, as_, ag)
    for name, ref in (("sw_tau", tau), ("ssa", ssa), ("g", gg)):
        assert rel_err(_f(ds, name, n), ref, floor=1e-6 * np.abs(ref).max()) < 1e-12, name
    toa = gs["toa_src"] * 0.98
    up, dn, dr = O.rte_sw(tau, ssa, gg, atm.mu0, toa, atm.sfc_alb, atm.sfc_alb, bl, atm.top_at_1, do_broadband=True)
    for name, ref in (("sw_flux_up", up), ("sw_flux_dn", dn), ("sw_flux_dn_dir", dr)):
        assert rel_err(_f(ds, name, n), ref, floor=1e-6) < 1e-10, name


def test_driver_single_precision(case_dir, O, okd_lw, okd_sw, lut_lw, lut_sw):
    d, atm = case_dir
    ds = _run(d, "--cloud-optics", "--single-precision")
    assert ds.get_variable("sw_flux_up").dtype == np.float32
    n = atm.ncol
    rlw = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw))
    rsw = O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), tsi_scaling=np.full(n, 0.98))
    for k in ("flux_up", "flux_dn"):
        assert np.abs(_f(ds, "lw_" + k, n) - rlw[k]).max() < 0.05, k
        assert np.abs(_f(ds, "sw_" + k, n) - rsw[k]).max() < 0.05, k


def test_driver_rejects_unknown_option(case_dir):
    from rrtmgp_amd import driver
    d, _ = case_dir
    assert driver.main(["--no-such-option"]) == 1
