"""Writes the golden fixtures of tests/golden/ from the FP64 oracle (run from the repository root:
python tests/golden/make_golden.py).

Two cases of the config-0 recipe at 8 columns x 60 layers (the first 8 columns of the seeded
128-column workloads): c0_clear (clear sky, seed of config 0) and c1_allsky (clouds, seed of
config 1).  Each fixture directory holds two bundles (the reference's NetCDF names, see
rrtmgp_amd/coefficients.py): case/ (the inputs, as rte_rrtmgp_input.nc) and expected/ (the
oracle's outputs), plus tables.json with SHA-256 digests of the seeded k-distribution and cloud
tables, which are regenerated rather than stored (about 25 MB).

expected/ holds: the interpolation indices jtemp, jpress, tropo and jeta; tau, the three LW
source arrays and the SW tau/ssa at every 8th g-point (gpt_sub = 1, 9, 17, ...); sfc_src and
toa_src at those g-points; broadband and by-band fluxes (LW up/down/net, SW up/down/direct/net)
and heating rates [K/day].  These are the oracle's values, so the fixtures pin the oracle
against later change and give the HIP path a fixed target (tests/test_golden.py).
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "rte-rrtmgp-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from rrtmgp_amd import coefficients as co  # noqa: E402
from rrtmgp_amd import synthetic as syn  # noqa: E402

AVAILABLE = ["h2o", "co2", "o3", "n2o", "ch4"]
HERE = os.path.dirname(os.path.abspath(__file__))
NCOL, NLAY, GSTEP = 8, 60, 8


def table_digests():
    out = {}
    for name, obj in (("kdist_lw", syn.make_kdist(True, syn.SEED_KDIST_LW)),
                      ("kdist_sw", syn.make_kdist(False, syn.SEED_KDIST_SW)),
                      ("cloud_lw", syn.make_cloud_lut(True, syn.SEED_CLOUD_LW)),
                      ("cloud_sw", syn.make_cloud_lut(False, syn.SEED_CLOUD_SW))):
        h = hashlib.sha256()
        for f in obj.__dataclass_fields__:
            v = getattr(obj, f)
            if isinstance(v, np.ndarray):
                h.update(f.encode())
                h.update(np.ascontiguousarray(v, dtype=v.dtype).tobytes(order="F"))
        out[name] = h.hexdigest()
    return out


def case_atmosphere(name):
    if name == "c0_clear":
        full = syn.make_atmosphere(syn.SEED_ATMOS[0], 128, NLAY)
    else:
        full = syn.make_atmosphere(syn.SEED_ATMOS[1], 128, NLAY, cloudy=True)
    return syn.subset_columns(full, np.arange(NCOL))


def expected(name, atm):
    import oracle as O
    kl, ks = syn.make_kdist(True, syn.SEED_KDIST_LW), syn.make_kdist(False, syn.SEED_KDIST_SW)
    okl, oks = O.Kdist(kl, AVAILABLE), O.Kdist(ks, AVAILABLE)
    cloudy = atm.lwp is not None
    cl = O.Cloud(syn.make_cloud_lut(True, syn.SEED_CLOUD_LW)) if cloudy else None
    cs = O.Cloud(syn.make_cloud_lut(False, syn.SEED_CLOUD_SW)) if cloudy else None
    tsi = np.full(atm.ncol, atm.tsi / oks.scalar("tsi"))
    s = co.Schema()
    s.dim("x", atm.ncol)
    s.dim("lay", atm.nlay)
    s.dim("lev", atm.nlay + 1)
    col_dry = O.get_col_dry(atm.vmr["h2o"], atm.p_lev)
    go = O.gas_optics(okl, atm, col_dry=col_dry, interp=True)
    gsw = O.gas_optics(oks, atm, col_dry=col_dry)
    sub = lambda ng: np.arange(0, ng, GSTEP)
    s.put_fortran("jtemp", ("lay", "x"), go["jtemp"].astype(np.int32))
    s.put_fortran("jpress", ("lay", "x"), go["jpress"].astype(np.int32))
    s.put_fortran("tropo", ("lay", "x"), go["tropo"].astype(np.int32))
    s.put_fortran("jeta_lw", ("flavor_lw", "lay", "x", "pair"), go["jeta"].astype(np.int32))
    s.put_fortran("lw_tau", ("gpt_sub_lw", "lay", "x"), go["tau"][:, :, sub(okl.ngpt)])
    for k in ("lay_src", "lev_src_inc", "lev_src_dec"):
        s.put_fortran(k, ("gpt_sub_lw", "lay", "x"), go[k][:, :, sub(okl.ngpt)])
    s.put_fortran("sfc_src", ("gpt_sub_lw", "x"), go["sfc_src"][:, sub(okl.ngpt)])
    s.put_fortran("sw_tau", ("gpt_sub_sw", "lay", "x"), gsw["tau"][:, :, sub(oks.ngpt)])
    s.put_fortran("sw_ssa", ("gpt_sub_sw", "lay", "x"), gsw["ssa"][:, :, sub(oks.ngpt)])
    s.put_fortran("toa_src", ("gpt_sub_sw", "x"), gsw["toa_src"][:, sub(oks.ngpt)])
    lw = O.solve_lw(okl, atm, cl, bnd=True)
    sw = O.solve_sw(oks, atm, cs, bnd=True, tsi_scaling=tsi)
    for k in ("flux_up", "flux_dn", "flux_net"):
        s.put_fortran("lw_" + k, ("lev", "x"), lw[k])
    for k in ("bnd_flux_up", "bnd_flux_dn"):
        s.put_fortran("lw_" + k, ("band_lw", "lev", "x"), lw[k])
    for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net"):
        s.put_fortran("sw_" + k, ("lev", "x"), sw[k])
    for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_dn_dir"):
        s.put_fortran("sw_" + k, ("band_sw", "lev", "x"), sw[k])
    s.put_fortran("lw_heating_rate", ("lay", "x"), O.heating_rate(lw["flux_dn"] - lw["flux_up"], atm.p_lev))
    s.put_fortran("sw_heating_rate", ("lay", "x"), O.heating_rate(sw["flux_dn"] - sw["flux_up"], atm.p_lev))
    s.put("tsi_scaling", (), np.float64(tsi[0]))
    return s


def main():
    digests = table_digests()
    for name in ("c0_clear", "c1_allsky"):
        atm = case_atmosphere(name)
        d = os.path.join(HERE, name)
        co.write_bundle(os.path.join(d, "case"), co.case_schema(atm))
        co.write_bundle(os.path.join(d, "expected"), expected(name, atm))
        with open(os.path.join(d, "tables.json"), "w") as f:
            json.dump(digests, f, indent=1)
        print("wrote", d)


if __name__ == "__main__":
    main()
