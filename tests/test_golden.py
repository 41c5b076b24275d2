"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the FP64 oracle).

CPU: the seeded tables still hash to the recorded digests, and the oracle reproduces every stored
value (integers exactly, reals within 1e-12 relative), so a later change of the oracle or of the
generator cannot pass unnoticed.  GPU: the HIP path meets the fixtures at the parity bars of the
rest of the suite: indices bit-exact, FP64 gas optics and fluxes within 1e-10 relative (1 W m-2
floor for fluxes), FP64 heating rates within 1e-8 K/day absolute plus 1e-8 relative, FP32 fluxes
within 0.05 W m-2.
"""
import json
import os

import numpy as np
import pytest

from conftest import AVAILABLE, rel_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("c0_clear", "c1_allsky")
GSTEP = 8


def _load(name):
    from rrtmgp_amd import coefficients as co
    d = os.path.join(HERE, name)
    case = co.read_bundle(os.path.join(d, "case"))
    exp = co.read_bundle(os.path.join(d, "expected"))
    with open(os.path.join(d, "tables.json")) as f:
        digests = json.load(f)
    return co.atmosphere_from_case(case), exp, digests


def test_table_digests_unchanged():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    now = mg.table_digests()
    for name in CASES:
        assert _load(name)[2] == now, f"the seeded tables changed; regenerate {name} deliberately"


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name, O, okd_lw, okd_sw, lut_lw, lut_sw):
    atm, e, _ = _load(name)
    cloudy = atm.lwp is not None
    col_dry = O.get_col_dry(atm.vmr["h2o"], atm.p_lev)
    go = O.gas_optics(okd_lw, atm, col_dry=col_dry, interp=True)
    for k in ("jtemp", "jpress", "tropo"):
        assert np.array_equal(go[k].astype(np.int32), e.get_fortran(k, np.int32)), k
    assert np.array_equal(go["jeta"].astype(np.int32), e.get_fortran("jeta_lw", np.int32))
    gl = slice(0, None, GSTEP)
    assert rel_err(go["tau"][:, :, gl], e.get_fortran("lw_tau"), floor=1e-300) < 1e-12
    assert rel_err(go["lay_src"][:, :, gl], e.get_fortran("lay_src"), floor=1e-300) < 1e-12
    lw = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw) if cloudy else None, bnd=True)
    sw = O.solve_sw(okd_sw, atm, O.Cloud(lut_sw) if cloudy else None, bnd=True,
                    tsi_scaling=np.full(atm.ncol, e.get_scalar("tsi_scaling")))
    for k in ("flux_up", "flux_dn", "flux_net", "bnd_flux_up", "bnd_flux_dn"):
        assert rel_err(lw[k], e.get_fortran("lw_" + k), floor=1.0) < 1e-12, k
    for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net", "bnd_flux_dn_dir"):
        assert rel_err(sw[k], e.get_fortran("sw_" + k), floor=1.0) < 1e-12, k


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_meets_fixture(name, kd_lw, kd_sw, lut_lw, lut_sw):
    from rrtmgp_amd import api
    atm, e, _ = _load(name)
    cloudy = atm.lwp is not None
    print(f"golden {name}: HIP FP64 and FP32 vs fixture", flush=True)
    for prec in (8, 4):
        kl, ks = api.KDist(kd_lw, AVAILABLE, prec), api.KDist(kd_sw, AVAILABLE, prec)
        if prec == 8:
            # gas optics at the stored g-points
            import oracle as O   # only for col_dry of the same inputs (pinned by its own tests)
            col_dry = O.get_col_dry(atm.vmr["h2o"], atm.p_lev)
            g = api.gas_optics(kl, atm, api.GasConcs(atm.vmr, 8), col_dry)
            gl = slice(0, None, GSTEP)
            assert rel_err(g["tau"].to_host()[:, :, gl], e.get_fortran("lw_tau"), floor=1e-300) < 1e-10
            for k in ("lay_src", "lev_src_inc", "lev_src_dec"):
                assert rel_err(g[k].to_host()[:, :, gl], e.get_fortran(k), floor=1e-300) < 1e-10, k
            assert rel_err(g["sfc_src"].to_host()[:, gl], e.get_fortran("sfc_src"), floor=1e-300) < 1e-10
            s = api.gas_optics(ks, atm, api.GasConcs(atm.vmr, 8), col_dry)
            assert rel_err(s["tau"].to_host()[:, :, gl], e.get_fortran("sw_tau"), floor=1e-300) < 1e-10
            assert np.abs(s["ssa"].to_host()[:, :, gl] - e.get_fortran("sw_ssa")).max() < 1e-12
        da = api.DeviceAtmosphere(atm, prec, cloudy=cloudy)
        lw = api.SolverLW(kl, api.Cloud(lut_lw, prec) if cloudy else None).solve(da, bnd=True)
        tsi = np.full(atm.ncol, e.get_scalar("tsi_scaling"))
        sw = api.SolverSW(ks, api.Cloud(lut_sw, prec) if cloudy else None).solve(da, tsi, bnd=True)
        pairs = [("lw_" + k, lw[k]) for k in ("flux_up", "flux_dn", "flux_net", "bnd_flux_up", "bnd_flux_dn")]
        pairs += [("sw_" + k, sw[k]) for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net", "bnd_flux_dn_dir")]
        for k, d in pairs:
            got, want = d.to_host().astype(np.float64), e.get_fortran(k)
            if prec == 8:
                assert rel_err(got, want, floor=1.0) < 1e-10, k
            else:
                assert np.abs(got - want).max() < 0.05, (k, np.abs(got - want).max())
        if prec == 8:
            # Heating rates differ by what the 1e-10 flux bar allows: c 2e-10 max|F| / |dp|.
            c = 9.80665 / 1004.64 * 86400.0
            dp = np.abs(np.diff(atm.p_lev, axis=1))
            for band, f in (("lw", lw), ("sw", sw)):
                hr = api.heating_rate(f["flux_up"], f["flux_dn"], atm.p_lev, 8).to_host()
                want = e.get_fortran(f"{band}_heating_rate")
                fmax = max(np.abs(e.get_fortran(f"{band}_flux_up")).max(), np.abs(e.get_fortran(f"{band}_flux_dn")).max())
                tol = c * 2e-10 * max(fmax, 1.0) / dp
                assert np.all(np.abs(hr - want) <= tol), (band, float((np.abs(hr - want) / tol).max()))
