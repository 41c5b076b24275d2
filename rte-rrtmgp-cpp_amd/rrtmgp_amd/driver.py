"""Stand-alone column-radiation driver (src_test/test_rte_rrtmgp.cpp:168-611).

Reads rte_rrtmgp_input.nc and the coefficient files coefficients_lw.nc, coefficients_sw.nc,
cloud_coefficients_lw.nc, cloud_coefficients_sw.nc and aerosol_optics.nc from the working
directory, solves the longwave and shortwave on the GPU through the drivers
(Radiation_solver_{longwave,shortwave}_gpu) and writes rte_rrtmgp_output.nc with the
reference's dimensions, variable names, dimension orders and types:

    python -m rrtmgp_amd.driver [--no-shortwave] [--cloud-optics] [--output-optical] ...

Options are the reference's (same names, defaults and --no- prefix; -h lists them), plus
--single-precision, which takes the place of the reference's compile-time RTE_USE_SP: the FP32
build (output variables float) instead of the FP64 one (double).

The output file is NetCDF-3 (64-bit offset, or 64-bit data when a variable passes 4 GiB) where
the reference writes NetCDF-4; the schema is the same.  With --output-optical the optical
properties are those after the cloud (and aerosol) additions, as the reference stores them
(Radiation_solver.cpp:488-503, 782-800), formed by a second pass of the class API's gas, cloud
and aerosol optics next to the drivers' flux computation.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

OPTIONS = {
    "shortwave": (True, "Enable computation of shortwave radiation."),
    "longwave": (True, "Enable computation of longwave radiation."),
    "fluxes": (True, "Enable computation of fluxes."),
    "cloud-optics": (False, "Enable cloud optics."),
    "aerosol-optics": (False, "Enable aerosol optics."),
    "output-optical": (False, "Enable output of optical properties."),
    "output-bnd-fluxes": (False, "Enable output of band fluxes."),
    "delta-cloud": (True, "delta-scaling of cloud optical properties"),
    "delta-aerosol": (False, "delta-scaling of aerosol optical properties"),
    "single-precision": (False, "FP32 build (the reference's RTE_USE_SP) instead of FP64."),
}

GASES = ("h2o", "co2", "o3", "n2o", "co", "ch4", "o2", "n2", "ccl4", "cfc11", "cfc12", "cfc22",
         "hfc143a", "hfc125", "hfc23", "hfc32", "hfc134a", "cf4", "no2")
AEROSOLS = tuple(f"aermr{i:02d}" for i in range(1, 12))


def message(s: str):
    print(s, flush=True)


def parse_options(argv) -> tuple[dict, bool]:
    """(switches, help_requested), as parse_command_line_options (test_rte_rrtmgp.cpp:108-153)."""
    sw = {k: v[0] for k, v in OPTIONS.items()}
    for raw in argv:
        arg = raw.strip()
        if arg in ("-h", "--help"):
            message("Possible usage:")
            for k, (_, text) in OPTIONS.items():
                message(f"{'--' + k:<30}{text}")
            return sw, True
        if not arg.startswith("--"):
            raise ValueError(f"{arg} is an illegal command line option.")
        name, enable = arg[2:], True
        if name.startswith("no-"):
            name, enable = name[3:], False
        if name not in sw:
            raise ValueError(f"{arg[2:]} is an illegal command line option.")
        sw[name] = enable
    return sw, False


class Case:
    """The input file's fields in the driver layout: (ncol, nz) Fortran order, ncol = x * y with x
    fastest, as test_rte_rrtmgp.cpp:205-293 reads them."""

    def __init__(self, ds, sw: dict):
        self.ds = ds
        self.nx, self.ny = ds.get_dimension_size("x"), ds.get_dimension_size("y")
        self.ncol = self.nx * self.ny
        self.nlay, self.nlev = ds.get_dimension_size("lay"), ds.get_dimension_size("lev")
        f = self.field
        self.p_lay, self.t_lay = f("p_lay", self.nlay), f("t_lay", self.nlay)
        self.p_lev, self.t_lev = f("p_lev", self.nlev), f("t_lev", self.nlev)
        self.col_dry = f("col_dry", self.nlay) if "col_dry" in ds.variables else None
        self.vmr = {}
        for g in GASES:
            v = self.gas(g)
            if v is not None:
                self.vmr[g] = v
        self.lwp = self.iwp = self.rel = self.rei = None
        if sw["cloud-optics"]:
            self.lwp, self.iwp, self.rel, self.rei = (f(k, self.nlay) for k in ("lwp", "iwp", "rel", "rei"))
        self.rh, self.aerosols = None, None
        if sw["aerosol-optics"]:
            self.rh = f("rh", self.nlay)
            self.aerosols = {}
            for a in AEROSOLS:
                if a not in ds.variables:
                    raise ValueError(f'Aerosol type "{a}" not available in input file.')
                self.aerosols[a] = self.profile_or_field(a)
        self.top_at_1 = bool(self.p_lay[0, 0] < self.p_lay[0, -1])

    def field(self, name, nz):
        a = self.ds.get_variable(name, (nz, self.ny, self.nx))
        return np.asfortranarray(a.reshape(nz, self.ncol).T.astype(np.float64))

    def surface(self, name, nbnd=None):
        if nbnd is None:
            return self.ds.get_variable(name, (self.ny, self.nx)).reshape(self.ncol).astype(np.float64)
        a = self.ds.get_variable(name, (self.ny, self.nx, nbnd))
        return np.asfortranarray(a.reshape(self.ncol, nbnd).T.astype(np.float64))

    def profile_or_field(self, name):
        v = self.ds.variables[name]
        if len(v.dimensions) == 1:
            if v.dimensions[0] != "lay" or v.shape[0] != self.nlay:
                raise ValueError(f'Illegal dimensions of "{name}" in input')
            prof = self.ds.get_variable(name).astype(np.float64)
            return np.asfortranarray(np.repeat(prof[None, :], self.ncol, axis=0))
        if len(v.dimensions) == 3 and tuple(v.dimensions) == ("lay", "y", "x"):
            return self.field(name, self.nlay)
        raise ValueError(f'Illegal dimensions of "{name}" in input')

    def gas(self, g):
        """read_and_set_vmr (test_rte_rrtmgp.cpp:31-67): scalar, (lay) profile or (lay, y, x)."""
        name = "vmr_" + g
        if name not in self.ds.variables:
            message(f'WARNING: Gas "{g}" not available in input file.')
            return None
        if len(self.ds.variables[name].dimensions) == 0:
            return np.asfortranarray([[float(self.ds.get_variable(name))]])
        return self.profile_or_field(name)


def _c_order(a, nx, ny):
    """A Fortran (ncol, nz[, ng]) array as the C-ordered ([ng,] nz, y, x) output variable."""
    a = np.asarray(a)
    t = a.T
    return np.ascontiguousarray(t.reshape(t.shape[:-1] + (ny, nx)))


def _load_kdist(path, available, prec):
    from . import api, coefficients, netcdf
    return api.KDist(coefficients.kdist_from_schema(netcdf.to_schema(netcdf.Dataset(path))), available, prec)


def _load_cloud(path, prec):
    from . import api, coefficients, netcdf
    return api.Cloud(coefficients.cloud_from_schema(netcdf.to_schema(netcdf.Dataset(path))), prec)


def _host(d, dt):
    return d.to_host().astype(dt)


def solve_radiation(argv=None, workdir: str = ".") -> int:
    from . import api, coefficients, device, netcdf
    from .synthetic import Atmosphere
    argv = sys.argv[1:] if argv is None else argv
    message("###### Starting RTE+RRTMGP solver ######")
    sw, stop = parse_options(argv)
    if stop:
        return 0
    message("Solver settings:")
    for k in OPTIONS:
        message(f"{k:<20} = {str(sw[k]).lower()}")
    prec = 4 if sw["single-precision"] else 8
    dt = np.float32 if prec == 4 else np.float64
    path = lambda name: os.path.join(workdir, name)       # noqa: E731

    message("Reading atmospheric input data from NetCDF.")
    case = Case(netcdf.Dataset(path("rte_rrtmgp_input.nc")), sw)
    nx, ny, ncol, nlay, nlev = case.nx, case.ny, case.ncol, case.nlay, case.nlev
    available = list(case.vmr)

    message("Preparing NetCDF output file.")
    out = netcdf.Writer(path("rte_rrtmgp_output.nc"))
    for name, size in (("x", nx), ("y", ny), ("lay", nlay), ("lev", nlev), ("pair", 2)):
        out.add_dimension(name, size)
    out.add_variable("p_lay", ("lay", "y", "x"), _c_order(case.p_lay, nx, ny).astype(dt))
    out.add_variable("p_lev", ("lev", "y", "x"), _c_order(case.p_lev, nx, ny).astype(dt))

    def atmosphere(emis=None, alb=None, mu0=None):
        return Atmosphere(p_lay=case.p_lay, p_lev=case.p_lev, t_lay=case.t_lay, t_lev=case.t_lev,
                          t_sfc=case.surface("t_sfc") if "t_sfc" in case.ds.variables else np.zeros(ncol),
                          vmr=case.vmr, emis_sfc=emis, sfc_alb=alb, mu0=mu0, tsi=0.0, lwp=case.lwp,
                          iwp=case.iwp, rel=case.rel, rei=case.rei, top_at_1=case.top_at_1)

    if sw["longwave"]:
        message("Initializing the longwave solver.")
        kd = _load_kdist(path("coefficients_lw.nc"), available, prec)
        cloud = _load_cloud(path("cloud_coefficients_lw.nc"), prec) if sw["cloud-optics"] else None
        nbnd, ngpt = kd.nbnd, kd.ngpt
        emis = case.surface("emis_sfc", nbnd)
        atm = atmosphere(emis=emis, alb=np.zeros((1, ncol)), mu0=np.ones(ncol))
        da = api.DeviceAtmosphere(atm, prec)
        message("Solving the longwave radiation.")
        t0 = time.perf_counter()
        fl = None
        if sw["fluxes"]:
            fl = api.SolverLW(kd, cloud).solve(da, bnd=sw["output-bnd-fluxes"], col_dry=case.col_dry)
        opt = None
        if sw["output-optical"]:
            gc = da.gas_concs
            cd = case.col_dry if case.col_dry is not None else api.get_col_dry(case.vmr["h2o"], case.p_lev, prec)
            opt = api.gas_optics(kd, atm, gc, cd)
            if cloud is not None:
                ct = cloud.optics(case.lwp, case.iwp, case.rel, case.rei, False)
                api.add_to(kd, opt["tau"], ct)
        device.synchronize()
        message(f"Duration longwave solver: {1e3 * (time.perf_counter() - t0):f} (ms)")
        message("Storing the longwave output.")
        out.add_dimension("gpt_lw", ngpt)
        out.add_dimension("band_lw", nbnd)
        out.add_variable("lw_band_lims_wvn", ("band_lw", "pair"),
                         np.ascontiguousarray(np.asarray(kd.src.band_lims_wavenum).T).astype(dt))
        if opt is not None:
            out.add_variable("lw_band_lims_gpt", ("band_lw", "pair"),
                             np.ascontiguousarray(np.asarray(kd.band_lims).T).astype(np.int32))
            g4 = ("gpt_lw", "lay", "y", "x")
            out.add_variable("lw_tau", g4, _c_order(_host(opt["tau"], dt), nx, ny))
            out.add_variable("lay_source", g4, _c_order(_host(opt["lay_src"], dt), nx, ny))
            out.add_variable("lev_source_inc", g4, _c_order(_host(opt["lev_src_inc"], dt), nx, ny))
            out.add_variable("lev_source_dec", g4, _c_order(_host(opt["lev_src_dec"], dt), nx, ny))
            out.add_variable("sfc_source", ("gpt_lw", "y", "x"), _c_order(_host(opt["sfc_src"], dt), nx, ny))
        if fl is not None:
            for k in ("up", "dn", "net"):
                out.add_variable(f"lw_flux_{k}", ("lev", "y", "x"), _c_order(_host(fl[f"flux_{k}"], dt), nx, ny))
            if sw["output-bnd-fluxes"]:
                for k in ("up", "dn", "net"):
                    out.add_variable(f"lw_bnd_flux_{k}", ("band_lw", "lev", "y", "x"),
                                     _c_order(_host(fl[f"bnd_flux_{k}"], dt), nx, ny))
        del da, fl, opt

    if sw["shortwave"]:
        message("Initializing the shortwave solver.")
        kd = _load_kdist(path("coefficients_sw.nc"), available, prec)
        cloud = _load_cloud(path("cloud_coefficients_sw.nc"), prec) if sw["cloud-optics"] else None
        aerosol = None
        if sw["aerosol-optics"]:
            lut = coefficients.aerosol_lut(netcdf.Dataset(path("aerosol_optics.nc")))
            aerosol = api.Aerosol(lut, prec)
        nbnd, ngpt = kd.nbnd, kd.ngpt
        mu0 = case.surface("mu0")
        alb_dir, alb_dif = case.surface("sfc_alb_dir", nbnd), case.surface("sfc_alb_dif", nbnd)
        if "tsi" in case.ds.variables:
            tsi_scaling = case.surface("tsi") / kd.scalar("tsi")
        elif "tsi_scaling" in case.ds.variables:
            tsi_scaling = np.full(ncol, float(case.ds.get_variable("tsi_scaling")))
        else:
            tsi_scaling = np.ones(ncol)
        atm = atmosphere(emis=np.zeros((1, ncol)), alb=alb_dir, mu0=mu0)
        da = api.DeviceAtmosphere(atm, prec)
        aer = None
        if aerosol is not None:
            aer = (case.rh, api.GasConcs(case.aerosols, prec))
        message("Solving the shortwave radiation.")
        t0 = time.perf_counter()
        fl = None
        if sw["fluxes"]:
            fl = api.SolverSW(kd, cloud, aerosol).solve(
                da, tsi_scaling, bnd=sw["output-bnd-fluxes"], delta_cloud=sw["delta-cloud"], col_dry=case.col_dry,
                aerosol=aer, delta_aerosol=sw["delta-aerosol"], sfc_alb_dif=alb_dif)
        opt = None
        if sw["output-optical"]:
            cd = case.col_dry if case.col_dry is not None else api.get_col_dry(case.vmr["h2o"], case.p_lev, prec)
            opt = api.gas_optics(kd, atm, da.gas_concs, cd)
            if cloud is not None:
                ct, cs, cg = cloud.optics(case.lwp, case.iwp, case.rel, case.rei, True,
                                          delta_scale=sw["delta-cloud"])
                api.add_to(kd, opt["tau"], ct, opt["ssa"], opt["g"], cs, cg)
            if aerosol is not None:
                at, as_, ag = aerosol.optics(aer[1], case.rh, case.p_lev, delta_scale=sw["delta-aerosol"])
                api.add_to(kd, opt["tau"], at, opt["ssa"], opt["g"], as_, ag)
        device.synchronize()
        message(f"Duration shortwave solver: {1e3 * (time.perf_counter() - t0):f} (ms)")
        message("Storing the shortwave output.")
        out.add_dimension("gpt_sw", ngpt)
        out.add_dimension("band_sw", nbnd)
        out.add_variable("sw_band_lims_wvn", ("band_sw", "pair"),
                         np.ascontiguousarray(np.asarray(kd.src.band_lims_wavenum).T).astype(dt))
        if opt is not None:
            out.add_variable("sw_band_lims_gpt", ("band_sw", "pair"),
                             np.ascontiguousarray(np.asarray(kd.band_lims).T).astype(np.int32))
            g4 = ("gpt_sw", "lay", "y", "x")
            out.add_variable("sw_tau", g4, _c_order(_host(opt["tau"], dt), nx, ny))
            out.add_variable("ssa", g4, _c_order(_host(opt["ssa"], dt), nx, ny))
            out.add_variable("g", g4, _c_order(_host(opt["g"], dt), nx, ny))
            toa = opt["toa_src"].to_host().astype(np.float64) * tsi_scaling[:, None]
            out.add_variable("toa_source", ("gpt_sw", "y", "x"), _c_order(toa.astype(dt), nx, ny))
        if fl is not None:
            for k in ("up", "dn", "dn_dir", "net"):
                out.add_variable(f"sw_flux_{k}", ("lev", "y", "x"), _c_order(_host(fl[f"flux_{k}"], dt), nx, ny))
            if sw["output-bnd-fluxes"]:
                for k in ("up", "dn", "dn_dir", "net"):
                    out.add_variable(f"sw_bnd_flux_{k}", ("band_sw", "lev", "y", "x"),
                                     _c_order(_host(fl[f"bnd_flux_{k}"], dt), nx, ny))
        del da, fl, opt

    out.close()
    message("###### Finished RTE+RRTMGP solver ######")
    return 0


def main(argv=None) -> int:
    try:
        return solve_radiation(argv)
    except Exception as e:          # the reference's main: report and return 1
        message(f"EXCEPTION: {e}")
        return 1


if __name__ == "__main__":
    sys.exit(main())
