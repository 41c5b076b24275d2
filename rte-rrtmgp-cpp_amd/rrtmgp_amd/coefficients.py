"""RRTMGP coefficient sets by variable name: the reference's NetCDF schema, and the build's bundle format.

The reference reads its k-distributions and cloud look-up tables from NetCDF files with fixed
variable names and dimension orders (src_test/Radiation_solver.cpp:70-284 for
rrtmgp-gas-{lw,sw}-*.nc, :286-327 for the cloud tables, :329-366 for aerosol_optics.nc).  Each
variable is stored in C order over its NetCDF dimensions, so its bytes are the reference's
Fortran-order Array with the dimension list reversed.  This module maps between that schema and
the in-memory `synthetic.KDist` / `synthetic.CloudLUT`, and stores a schema as a *bundle*:

    <dir>/manifest.txt   one line per dimension ("dim <name> <size>") and per variable
                         ("var <name> <type> <ndim> <dim names...>"), type in f8 f4 i4 i1 S1
    <dir>/<name>.bin     the variable's values, little-endian, C order over its dimensions

A bundle is what the C++ loaders of include/Coefficient_bundle.h read (load_and_init_gas_optics
and load_and_init_cloud_optics, the counterparts of the reference's NetCDF loaders); netcdf.py
converts a NetCDF coefficient file into one.
"""
from __future__ import annotations

import os

import numpy as np

from . import synthetic as syn

_TYPES = {"f8": np.dtype("<f8"), "f4": np.dtype("<f4"), "i4": np.dtype("<i4"), "i1": np.dtype("i1"),
          "S1": np.dtype("S1")}


def _type_code(a: np.ndarray) -> str:
    for code, dt in _TYPES.items():
        if a.dtype.kind == dt.kind and a.dtype.itemsize == dt.itemsize:
            return code
    raise TypeError(f"unsupported dtype {a.dtype}")


class Schema:
    """Dimensions (name -> size) and variables (name -> (dimension names, C-order array))."""

    def __init__(self):
        self.dims: dict[str, int] = {}
        self.vars: dict[str, tuple[tuple[str, ...], np.ndarray]] = {}

    def dim(self, name: str, size: int):
        size = int(size)
        if self.dims.setdefault(name, size) != size:
            raise ValueError(f"dimension {name}: {self.dims[name]} != {size}")

    def put(self, name: str, dims: tuple, value):
        a = np.array(value, order="C")
        if a.ndim != len(dims):
            raise ValueError(f"{name}: {a.ndim} dims given for {len(dims)} names")
        for d, n in zip(dims, a.shape):
            self.dim(d, n)
        self.vars[name] = (tuple(dims), a)

    def put_fortran(self, name: str, dims: tuple, fortran_array):
        """A Fortran-order array of dims reversed(dims) (the reference's Array layout)."""
        a = np.asarray(fortran_array)
        self.put(name, dims, np.ascontiguousarray(a.transpose(*reversed(range(a.ndim)))))

    def put_strings(self, name: str, dim: str, strings, string_len: int = 32):
        self.dim("string_len", string_len)
        a = np.full((len(strings), string_len), b" ", dtype="S1")
        for i, s in enumerate(strings):
            b = s.encode()
            if len(b) > string_len:
                raise ValueError(f"{name}: '{s}' longer than {string_len}")
            a[i, :len(b)] = np.frombuffer(b, dtype="S1")
        self.put(name, (dim, "string_len"), a)

    def has(self, name: str) -> bool:
        return name in self.vars

    def get(self, name: str) -> np.ndarray:
        return self.vars[name][1]

    def get_fortran(self, name: str, dtype=np.float64) -> np.ndarray:
        """The variable as a Fortran-order array of the reversed dims (the reference's Array)."""
        a = self.get(name)
        return np.asfortranarray(a.transpose(*reversed(range(a.ndim))).astype(dtype))

    def get_scalar(self, name: str) -> float:
        return float(np.asarray(self.get(name)).reshape(-1)[0])

    def get_strings(self, name: str) -> list:
        a = self.get(name)
        return [b"".join(row.tolist()).decode().strip(" \x00") for row in a.reshape(a.shape[0], -1)]


# ---------------------------------------------------------------------------------------------
# k-distributions (src_test/Radiation_solver.cpp:70-284: names and dimension orders)

def kdist_schema(kd: syn.KDist) -> Schema:
    s = Schema()
    s.put_strings("gas_names", "absorber", kd.gas_names)
    s.put_fortran("key_species", ("bnd", "atmos_layer", "pair"), kd.key_species.astype(np.int32))
    s.put_fortran("bnd_limits_wavenumber", ("bnd", "pair"), kd.band_lims_wavenum)
    s.put_fortran("bnd_limits_gpt", ("bnd", "pair"), kd.band2gpt.astype(np.int32))
    s.put("press_ref", ("pressure",), kd.press_ref)
    s.put("temp_ref", ("temperature",), kd.temp_ref)
    s.put("absorption_coefficient_ref_P", (), np.float64(kd.temp_ref_p))
    s.put("absorption_coefficient_ref_T", (), np.float64(kd.temp_ref_t))
    s.put("press_ref_trop", (), np.float64(kd.press_ref_trop))
    s.put_strings("gas_minor", "minor_absorber", kd.gas_minor)
    s.put_strings("identifier_minor", "minor_absorber", kd.identifier_minor)
    for side in ("lower", "upper"):
        iv = f"minor_absorber_intervals_{side}"
        s.put_fortran(f"kminor_{side}", ("temperature", "mixing_fraction", f"contributors_{side}"),
                      getattr(kd, f"kminor_{side}"))
        s.put_strings(f"minor_gases_{side}", iv, getattr(kd, f"minor_gases_{side}"))
        s.put_fortran(f"minor_limits_gpt_{side}", (iv, "pair"),
                      getattr(kd, f"minor_limits_gpt_{side}").astype(np.int32))
        for v in ("minor_scales_with_density", "scale_by_complement"):
            s.put(f"{v}_{side}", (iv,), getattr(kd, f"{v}_{side}").astype(np.int32))
        s.put_strings(f"scaling_gas_{side}", iv, getattr(kd, f"scaling_gas_{side}"))
        s.put(f"kminor_start_{side}", (iv,), getattr(kd, f"kminor_start_{side}").astype(np.int32))
    s.put_fortran("vmr_ref", ("temperature", "absorber_ext", "atmos_layer"), kd.vmr_ref)
    s.put_fortran("kmajor", ("temperature", "pressure_interp", "mixing_fraction", "gpt"), kd.kmajor)
    if kd.rayl_lower is not None:
        s.put_fortran("rayl_lower", ("temperature", "mixing_fraction", "gpt"), kd.rayl_lower)
        s.put_fortran("rayl_upper", ("temperature", "mixing_fraction", "gpt"), kd.rayl_upper)
    if kd.totplnk is not None:
        s.put_fortran("totplnk", ("bnd", "temperature_Planck"), kd.totplnk)
        s.put_fortran("plank_fraction", ("temperature", "pressure_interp", "mixing_fraction", "gpt"),
                      kd.planck_frac)
    else:
        for v in ("quiet", "facular", "sunspot"):
            s.put(f"solar_source_{v}", ("gpt",), getattr(kd, f"solar_src_{v}"))
        s.put("tsi_default", (), np.float64(kd.tsi_default))
        s.put("mg_default", (), np.float64(kd.mg_default))
        s.put("sb_default", (), np.float64(kd.sb_default))
    return s


def kdist_from_schema(s: Schema) -> syn.KDist:
    """The k-distribution of a coefficient file's variables (the reference's reader, in numpy)."""
    i32 = np.int32
    kw = dict(
        gas_names=s.get_strings("gas_names"),
        key_species=s.get_fortran("key_species", i32),
        band2gpt=s.get_fortran("bnd_limits_gpt", i32),
        band_lims_wavenum=s.get_fortran("bnd_limits_wavenumber"),
        press_ref=s.get_fortran("press_ref"), press_ref_trop=s.get_scalar("press_ref_trop"),
        temp_ref=s.get_fortran("temp_ref"), temp_ref_p=s.get_scalar("absorption_coefficient_ref_P"),
        temp_ref_t=s.get_scalar("absorption_coefficient_ref_T"), vmr_ref=s.get_fortran("vmr_ref"),
        kmajor=s.get_fortran("kmajor"), gas_minor=s.get_strings("gas_minor"),
        identifier_minor=s.get_strings("identifier_minor"))
    for side in ("lower", "upper"):
        kw[f"kminor_{side}"] = s.get_fortran(f"kminor_{side}")
        kw[f"minor_gases_{side}"] = s.get_strings(f"minor_gases_{side}")
        kw[f"minor_limits_gpt_{side}"] = s.get_fortran(f"minor_limits_gpt_{side}", i32)
        kw[f"minor_scales_with_density_{side}"] = s.get_fortran(f"minor_scales_with_density_{side}", np.int8)
        kw[f"scale_by_complement_{side}"] = s.get_fortran(f"scale_by_complement_{side}", np.int8)
        kw[f"scaling_gas_{side}"] = s.get_strings(f"scaling_gas_{side}")
        kw[f"kminor_start_{side}"] = s.get_fortran(f"kminor_start_{side}", i32)
    is_lw = s.has("totplnk")
    kd = syn.KDist(is_lw=is_lw, **kw)
    if s.has("rayl_lower"):
        kd.rayl_lower = s.get_fortran("rayl_lower")
        kd.rayl_upper = s.get_fortran("rayl_upper")
    if is_lw:
        kd.totplnk = s.get_fortran("totplnk")
        kd.planck_frac = s.get_fortran("plank_fraction")
    else:
        for v in ("quiet", "facular", "sunspot"):
            setattr(kd, f"solar_src_{v}", s.get_fortran(f"solar_source_{v}"))
        kd.tsi_default = s.get_scalar("tsi_default")
        kd.mg_default = s.get_scalar("mg_default")
        kd.sb_default = s.get_scalar("sb_default")
    return kd


# ---------------------------------------------------------------------------------------------
# cloud look-up tables (src_test/Radiation_solver.cpp:286-327)

def cloud_schema(c: syn.CloudLUT) -> Schema:
    s = Schema()
    s.put_fortran("bnd_limits_wavenumber", ("nband", "pair"), c.band_lims_wvn)
    for v in ("radliq_lwr", "radliq_upr", "radliq_fac", "radice_lwr", "radice_upr", "radice_fac"):
        s.put(v, (), np.float64(getattr(c, v)))
    for v in ("extliq", "ssaliq", "asyliq"):
        s.put_fortran(f"lut_{v}", ("nband", "nsize_liq"), getattr(c, f"lut_{v}"))
    for v in ("extice", "ssaice", "asyice"):
        s.put_fortran(f"lut_{v}", ("nrghice", "nband", "nsize_ice"), getattr(c, f"lut_{v}"))
    return s


def cloud_from_schema(s: Schema) -> syn.CloudLUT:
    kw = {v: s.get_scalar(v) for v in ("radliq_lwr", "radliq_upr", "radliq_fac",
                                        "radice_lwr", "radice_upr", "radice_fac")}
    for v in ("extliq", "ssaliq", "asyliq", "extice", "ssaice", "asyice"):
        kw[f"lut_{v}"] = s.get_fortran(f"lut_{v}")
    return syn.CloudLUT(band_lims_wvn=s.get_fortran("bnd_limits_wavenumber"), **kw)


# ---------------------------------------------------------------------------------------------
# aerosol look-up tables (src_test/Radiation_solver.cpp:329-366)

def aerosol_lut(ds) -> dict:
    """The tables load_and_init_aerosol_optics reads from aerosol_optics.nc, in the layouts the
    Aerosol_optics constructor takes: rh_upper (nhum) from relative_humidity2, mext/ssa/g_phobic
    (nbnd, nphobic) and mext/ssa/g_philic (nbnd, nhum, nphilic) from the *_sw_* variables.  ds is a
    netcdf.Dataset (or anything with get_dimension_size and get_variable(name, dims))."""
    n_band = ds.get_dimension_size("band_sw")
    n_hum = ds.get_dimension_size("relative_humidity")
    n_philic = ds.get_dimension_size("hydrophilic")
    n_phobic = ds.get_dimension_size("hydrophobic")
    out = {"rh_upper": np.asarray(ds.get_variable("relative_humidity2", (n_hum,)), dtype=np.float64)}
    for key, name in (("mext", "mass_ext"), ("ssa", "ssa"), ("g", "asymmetry")):
        phobic = ds.get_variable(f"{name}_sw_hydrophobic", (n_phobic, n_band))
        philic = ds.get_variable(f"{name}_sw_hydrophilic", (n_philic, n_hum, n_band))
        # C order over the NetCDF dimensions = Fortran order over the reversed ones.
        out[f"{key}_phobic"] = np.asfortranarray(np.asarray(phobic, dtype=np.float64).T)
        out[f"{key}_philic"] = np.asfortranarray(np.asarray(philic, dtype=np.float64).T)
    return out


# ---------------------------------------------------------------------------------------------
# bundles

def write_bundle(path: str, s: Schema):
    os.makedirs(path, exist_ok=True)
    lines = ["# rrtmgp coefficient bundle v1"]
    lines += [f"dim {n} {v}" for n, v in s.dims.items()]
    for name, (dims, a) in s.vars.items():
        code = _type_code(a)
        lines.append(f"var {name} {code} {len(dims)} {' '.join(dims)}".rstrip())
        np.array(a, dtype=_TYPES[code], order="C").tofile(os.path.join(path, name + ".bin"))
    with open(os.path.join(path, "manifest.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


def read_bundle(path: str) -> Schema:
    s = Schema()
    with open(os.path.join(path, "manifest.txt")) as f:
        rows = [ln.split() for ln in f if ln.strip() and not ln.startswith("#")]
    for r in rows:
        if r[0] == "dim":
            s.dim(r[1], int(r[2]))
    for r in rows:
        if r[0] != "var":
            continue
        name, code, nd = r[1], r[2], int(r[3])
        dims = tuple(r[4:4 + nd])
        shape = tuple(s.dims[d] for d in dims)
        a = np.fromfile(os.path.join(path, name + ".bin"), dtype=_TYPES[code])
        if a.size != int(np.prod(shape, dtype=np.int64)):
            raise ValueError(f"bundle {path}: {name} holds {a.size} values, expected {shape}")
        s.vars[name] = (dims, a.reshape(shape))
    return s


# ---------------------------------------------------------------------------------------------
# case files: rte_rrtmgp_input.nc (src_test/test_rte_rrtmgp.cpp:205-293, :352-356, :464-488)

_CASE_GASES = ("h2o", "co2", "o3", "n2o", "co", "ch4", "o2", "n2", "ccl4", "cfc11", "cfc12", "cfc22",
               "hfc143a", "hfc125", "hfc23", "hfc32", "hfc134a", "cf4", "no2")


def case_schema(atm: syn.Atmosphere, tsi_scaling: float | None = None) -> Schema:
    """An atmosphere as the reference's input file: columns on x (y = 1), fields (lay|lev, y, x),
    surface fields (y, x[, band])."""
    s = Schema()
    ncol, nlay = atm.ncol, atm.nlay

    def field(name, a, nz):
        s.put(name, (nz, "y", "x"), np.asarray(a).T.reshape(-1, 1, ncol))

    s.dim("x", ncol)
    s.dim("y", 1)
    field("p_lay", atm.p_lay, "lay")
    field("t_lay", atm.t_lay, "lay")
    field("p_lev", atm.p_lev, "lev")
    field("t_lev", atm.t_lev, "lev")
    for g, v in atm.vmr.items():
        v = np.asarray(v)
        if v.size == 1:
            s.put("vmr_" + g, (), np.float64(v.reshape(-1)[0]))
        else:
            field("vmr_" + g, v, "lay")
    if atm.lwp is not None:
        for k in ("lwp", "iwp", "rel", "rei"):
            field(k, getattr(atm, k), "lay")
    s.put("t_sfc", ("y", "x"), np.asarray(atm.t_sfc).reshape(1, ncol))
    s.put("emis_sfc", ("y", "x", "band_lw"), np.asarray(atm.emis_sfc).T.reshape(1, ncol, -1))
    s.put("mu0", ("y", "x"), np.asarray(atm.mu0).reshape(1, ncol))
    alb = np.asarray(atm.sfc_alb).T.reshape(1, ncol, -1)
    s.put("sfc_alb_dir", ("y", "x", "band_sw"), alb)
    s.put("sfc_alb_dif", ("y", "x", "band_sw"), alb)
    if tsi_scaling is not None:
        s.put("tsi_scaling", (), np.float64(tsi_scaling))
    return s


def atmosphere_from_case(s: Schema) -> syn.Atmosphere:
    """An input file's atmosphere in the driver layout ((ncol, nz) Fortran order, ncol = x*y)."""
    ncol = s.dims["x"] * s.dims.get("y", 1)

    def field(name):
        a = s.get(name)
        return np.asfortranarray(a.reshape(a.shape[0], ncol).T.astype(np.float64))

    vmr = {}
    for g in _CASE_GASES:
        name = "vmr_" + g
        if not s.has(name):
            continue
        dims = s.vars[name][0]
        if len(dims) == 0:
            vmr[g] = np.asfortranarray([[s.get_scalar(name)]])
        elif len(dims) == 1:
            vmr[g] = np.asfortranarray(np.repeat(s.get(name)[None, :].astype(np.float64), ncol, axis=0))
        else:
            vmr[g] = field(name)

    def sfc(name):
        return np.asfortranarray(s.get(name).reshape(ncol, -1).T.astype(np.float64))

    p_lay = field("p_lay")
    atm = syn.Atmosphere(
        p_lay=p_lay, p_lev=field("p_lev"), t_lay=field("t_lay"), t_lev=field("t_lev"),
        t_sfc=s.get("t_sfc").reshape(ncol).astype(np.float64), vmr=vmr,
        emis_sfc=sfc("emis_sfc") if s.has("emis_sfc") else None,
        sfc_alb=sfc("sfc_alb_dir") if s.has("sfc_alb_dir") else None,
        mu0=s.get("mu0").reshape(ncol).astype(np.float64) if s.has("mu0") else None,
        tsi=1360.0, top_at_1=bool(p_lay[0, 0] < p_lay[0, -1]))
    if s.has("lwp"):
        atm.lwp, atm.iwp, atm.rel, atm.rei = (field(k) for k in ("lwp", "iwp", "rel", "rei"))
    if s.has("sfc_alb_dif"):
        atm.extra["sfc_alb_dif"] = sfc("sfc_alb_dif")
    return atm
