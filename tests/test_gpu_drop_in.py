"""The drop-in boundary, tested as a reference user would meet it.

tests/native/drop_in_driver.cpp is written like the reference's driver (test_rte_rrtmgp.cpp reading
the case; Radiation_solver.cu:419-580 / :683-893 calling the class API per column block with
Array_gpu::subset, Fluxes_broadband_gpu::reduce and Array_gpu::dump).  It is compiled against this
build's headers by tests/native/build_native.py (run by __graft_entry__.build()) and reads its
inputs from coefficient and case bundles in the reference's NetCDF schema.  Its fluxes are checked
against the FP64 oracle: within 1e-10 relative (floor 1 W m-2) in the FP64 build and within
0.05 W m-2 in the FP32 build.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import AVAILABLE, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "native", "_build", "drop_in_driver_{}")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(tmp_path_factory, syn, kd_lw, kd_sw, lut_lw, lut_sw, okd_lw, okd_sw, O):
    from rrtmgp_amd import coefficients as co
    d = tmp_path_factory.mktemp("drop_in")
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 37, 60, cloudy=True)
    tsi_scaling = atm.tsi / okd_sw.scalar("tsi")
    co.write_bundle(str(d / "case"), co.case_schema(atm, tsi_scaling=tsi_scaling))
    co.write_bundle(str(d / "lw"), co.kdist_schema(kd_lw))
    co.write_bundle(str(d / "sw"), co.kdist_schema(kd_sw))
    co.write_bundle(str(d / "cloud_lw"), co.cloud_schema(lut_lw))
    co.write_bundle(str(d / "cloud_sw"), co.cloud_schema(lut_sw))
    want_lw = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw), bnd=True)
    want_sw = O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), tsi_scaling=np.full(atm.ncol, tsi_scaling))
    return d, atm, want_lw, want_sw


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_reference_shaped_driver(case, prec):
    d, atm, want_lw, want_sw = case
    exe = EXE.format(prec)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    out = d / f"out_{prec}"
    out.mkdir(exist_ok=True)
    # 37 columns in blocks of 16: two full blocks and a residual block of 5.
    print(f"drop_in_driver_{prec}: 37 columns x 60 layers, blocks of 16", flush=True)
    r = subprocess.run([exe, str(d / "case"), str(d / "lw"), str(d / "sw"), str(d / "cloud_lw"),
                        str(d / "cloud_sw"), str(out), "16"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dt = np.float64 if prec == "f64" else np.float32
    ncol, nlev = atm.ncol, atm.nlay + 1

    def load(name, *shape):
        return np.fromfile(out / f"{name}.bin", dtype=dt).reshape(shape, order="F").astype(np.float64)

    got = {
        "lw_flux_up": (load("lw_flux_up", ncol, nlev), want_lw["flux_up"]),
        "lw_flux_dn": (load("lw_flux_dn", ncol, nlev), want_lw["flux_dn"]),
        "lw_flux_net": (load("lw_flux_net", ncol, nlev), want_lw["flux_net"]),
        "lw_bnd_flux_up": (load("lw_bnd_flux_up", ncol, nlev, 16), want_lw["bnd_flux_up"]),
        "lw_bnd_flux_dn": (load("lw_bnd_flux_dn", ncol, nlev, 16), want_lw["bnd_flux_dn"]),
        "sw_flux_up": (load("sw_flux_up", ncol, nlev), want_sw["flux_up"]),
        "sw_flux_dn": (load("sw_flux_dn", ncol, nlev), want_sw["flux_dn"]),
        "sw_flux_dn_dir": (load("sw_flux_dn_dir", ncol, nlev), want_sw["flux_dn_dir"]),
        "sw_flux_net": (load("sw_flux_net", ncol, nlev), want_sw["flux_net"]),
    }
    for name, (g, w) in got.items():
        if prec == "f64":
            assert rel_err(g, w, floor=1.0) < 1e-10, name
        else:
            assert np.abs(g - w).max() < 0.05, (name, np.abs(g - w).max())
