"""ctypes binding of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker; the product never imports it.  Arrays follow the reference's Fortran layouts
(first index fastest); numpy arrays are passed in Fortran order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librrtmgp_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_int8)


class OracleKdistFile(C.Structure):
    _fields_ = [
        ("ngas", C.c_int), ("gas_names", C.c_char_p),
        ("nbnd", C.c_int), ("ngpt", C.c_int), ("ntemp", C.c_int), ("npres", C.c_int),
        ("neta", C.c_int), ("nplanck", C.c_int),
        ("key_species", _ip), ("band2gpt", _ip), ("band_lims_wavenum", _dp),
        ("press_ref", _dp), ("press_ref_trop", C.c_double), ("temp_ref", _dp),
        ("temp_ref_p", C.c_double), ("temp_ref_t", C.c_double), ("vmr_ref", _dp),
        ("kmajor", _dp),
        ("nminor_ident", C.c_int), ("gas_minor", C.c_char_p), ("identifier_minor", C.c_char_p),
        ("nminor_lower", C.c_int), ("ncontrib_lower", C.c_int), ("kminor_lower", _dp),
        ("minor_gases_lower", C.c_char_p), ("minor_limits_gpt_lower", _ip),
        ("minor_scales_with_density_lower", _bp), ("scaling_gas_lower", C.c_char_p),
        ("scale_by_complement_lower", _bp), ("kminor_start_lower", _ip),
        ("nminor_upper", C.c_int), ("ncontrib_upper", C.c_int), ("kminor_upper", _dp),
        ("minor_gases_upper", C.c_char_p), ("minor_limits_gpt_upper", _ip),
        ("minor_scales_with_density_upper", _bp), ("scaling_gas_upper", C.c_char_p),
        ("scale_by_complement_upper", _bp), ("kminor_start_upper", _ip),
        ("totplnk", _dp), ("planck_frac", _dp),
        ("solar_src_quiet", _dp), ("solar_src_facular", _dp), ("solar_src_sunspot", _dp),
        ("tsi_default", C.c_double), ("mg_default", C.c_double), ("sb_default", C.c_double),
        ("rayl_lower", _dp), ("rayl_upper", _dp),
        ("n_available", C.c_int), ("available_gases", C.c_char_p),
    ]


class OracleGasConcs(C.Structure):
    _fields_ = [("n", C.c_int), ("names", C.c_char_p), ("data", C.POINTER(_dp)),
                ("dims", _ip)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run oracle/build_oracle.py)")
        L = C.CDLL(LIB_PATH)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_kdist_create.restype = C.c_void_p
        L.oracle_kdist_create.argtypes = [C.POINTER(OracleKdistFile)]
        L.oracle_kdist_destroy.argtypes = [C.c_void_p]
        L.oracle_kdist_info.argtypes = [C.c_void_p, _ip]
        L.oracle_kdist_get_int.argtypes = [C.c_void_p, C.c_char_p, _ip]
        L.oracle_kdist_get_real.argtypes = [C.c_void_p, C.c_char_p, _dp]
        L.oracle_kdist_gas_names.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.oracle_cloud_create.restype = C.c_void_p
        L.oracle_cloud_create.argtypes = [C.c_int] + [C.c_double] * 4 + [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_cloud_destroy.argtypes = [C.c_void_p]
        for name in ("oracle_gas_optics", "oracle_cloud_optics", "oracle_solve_lw", "oracle_solve_sw",
                     "oracle_rte_lw", "oracle_rte_sw", "oracle_rte_lw_rescaled", "oracle_lw_solver_2stream",
                     "oracle_aerosol_optics"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


def _i8(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int8))


def _names(lst):
    return "\n".join(lst).encode()


def _check(rc):
    if rc != 0:
        msg = lib().oracle_last_error().decode()
        if rc == 2:
            raise ValueError(msg)
        raise RuntimeError(msg)


class Kdist:
    """Oracle k-distribution built from a synthetic.KDist (Gas_optics_rrtmgp ctor restated)."""

    def __init__(self, kd, available_gases):
        self.src = kd
        keep = []
        f = OracleKdistFile()
        arrays = {}

        def hold(name, arr, conv):
            a = conv(arr)
            keep.append(a)
            return a

        f.ngas = len(kd.gas_names)
        f.gas_names = _names(kd.gas_names)
        f.nbnd, f.ngpt = kd.nbnd, kd.ngpt
        f.ntemp = kd.temp_ref.size
        f.npres = kd.press_ref.size
        f.neta = kd.kmajor.shape[1]
        f.nplanck = kd.totplnk.shape[0] if kd.totplnk is not None else 0
        f.key_species = ptr(hold("ks", kd.key_species, _i32), C.c_int)
        f.band2gpt = ptr(hold("b2g", kd.band2gpt, _i32), C.c_int)
        f.band_lims_wavenum = ptr(hold("blw", kd.band_lims_wavenum, _f64))
        f.press_ref = ptr(hold("pr", kd.press_ref, _f64))
        f.press_ref_trop = kd.press_ref_trop
        f.temp_ref = ptr(hold("tr", kd.temp_ref, _f64))
        f.temp_ref_p, f.temp_ref_t = kd.temp_ref_p, kd.temp_ref_t
        f.vmr_ref = ptr(hold("vr", kd.vmr_ref, _f64))
        f.kmajor = ptr(hold("km", kd.kmajor, _f64))
        f.nminor_ident = len(kd.identifier_minor)
        f.gas_minor = _names(kd.gas_minor)
        f.identifier_minor = _names(kd.identifier_minor)
        for side in ("lower", "upper"):
            km = getattr(kd, f"kminor_{side}")
            setattr(f, f"nminor_{side}", len(getattr(kd, f"minor_gases_{side}")))
            setattr(f, f"ncontrib_{side}", km.shape[0])
            setattr(f, f"kminor_{side}", ptr(hold("", km, _f64)))
            setattr(f, f"minor_gases_{side}", _names(getattr(kd, f"minor_gases_{side}")))
            setattr(f, f"minor_limits_gpt_{side}", ptr(hold("", getattr(kd, f"minor_limits_gpt_{side}"), _i32), C.c_int))
            setattr(f, f"minor_scales_with_density_{side}", ptr(hold("", getattr(kd, f"minor_scales_with_density_{side}"), _i8), C.c_int8))
            setattr(f, f"scaling_gas_{side}", _names(getattr(kd, f"scaling_gas_{side}")))
            setattr(f, f"scale_by_complement_{side}", ptr(hold("", getattr(kd, f"scale_by_complement_{side}"), _i8), C.c_int8))
            setattr(f, f"kminor_start_{side}", ptr(hold("", getattr(kd, f"kminor_start_{side}"), _i32), C.c_int))
        if kd.totplnk is not None:
            f.totplnk = ptr(hold("tp", kd.totplnk, _f64))
            f.planck_frac = ptr(hold("pf", kd.planck_frac, _f64))
        if kd.solar_src_quiet is not None:
            f.solar_src_quiet = ptr(hold("", kd.solar_src_quiet, _f64))
            f.solar_src_facular = ptr(hold("", kd.solar_src_facular, _f64))
            f.solar_src_sunspot = ptr(hold("", kd.solar_src_sunspot, _f64))
            f.tsi_default, f.mg_default, f.sb_default = kd.tsi_default, kd.mg_default, kd.sb_default
            f.rayl_lower = ptr(hold("", kd.rayl_lower, _f64))
            f.rayl_upper = ptr(hold("", kd.rayl_upper, _f64))
        f.n_available = len(available_gases)
        f.available_gases = _names(list(available_gases))
        h = lib().oracle_kdist_create(C.byref(f))
        if not h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.h = C.c_void_p(h)
        info = np.zeros(14, dtype=np.int32)
        lib().oracle_kdist_info(self.h, ptr(info, C.c_int))
        (self.ngas, self.nflav, self.neta, self.npres, self.ntemp, self.ngpt, self.nbnd,
         self.nplanck, self.nminor_lower, self.nminor_upper, self.ncontrib_lower,
         self.ncontrib_upper, is_lw, self.idx_h2o) = [int(x) for x in info]
        self.is_lw = bool(is_lw)
        buf = C.create_string_buffer(4096)
        lib().oracle_kdist_gas_names(self.h, buf, 4096)
        self.gas_names = buf.value.decode().split("\n")
        self.band_lims = _i32(kd.band2gpt)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_kdist_destroy(self.h)
            self.h = None

    def get_int(self, name, n=100000):
        out = np.zeros(n, dtype=np.int32)
        k = lib().oracle_kdist_get_int(self.h, name.encode(), ptr(out, C.c_int))
        if k < 0:
            raise KeyError(name)
        return out[:k].copy()

    def get_real(self, name, n=4_000_000):
        out = np.zeros(n)
        k = lib().oracle_kdist_get_real(self.h, name.encode(), ptr(out))
        if k < 0:
            raise KeyError(name)
        return out[:k].copy()

    def scalar(self, name):
        return float(self.get_real(name, 1)[0])


class GasConcs:
    """Name -> vmr map marshalled for the oracle."""

    def __init__(self, vmr: dict):
        self.names = list(vmr.keys())
        self.arrays = [_f64(vmr[k]) for k in self.names]
        self.dims = _i32([d for a in self.arrays for d in a.shape])
        self.ptrs = (_dp * len(self.arrays))(*[ptr(a) for a in self.arrays])
        self.s = OracleGasConcs(len(self.names), _names(self.names), self.ptrs, ptr(self.dims, C.c_int))


def get_col_dry(vmr_h2o, plev):
    vmr_h2o, plev = _f64(vmr_h2o), _f64(plev)
    ncol, nlay = vmr_h2o.shape
    out = np.zeros((ncol, nlay), order="F")
    lib().oracle_get_col_dry(ncol, nlay, ptr(vmr_h2o), ptr(plev), ptr(out))
    return out


def gas_optics(kd: Kdist, atm, col_dry=None, interp=False):
    """Full gas optics of the oracle; returns a dict of outputs in Fortran layouts."""
    ncol, nlay = atm.p_lay.shape
    ngpt = kd.ngpt
    gc = GasConcs(atm.vmr)
    if col_dry is None:
        col_dry = get_col_dry(np.broadcast_to(atm.vmr["h2o"], (ncol, nlay)), atm.p_lev)
    col_dry = _f64(col_dry)
    o = {"tau": np.zeros((ncol, nlay, ngpt), order="F")}
    if kd.is_lw:
        for k in ("lay_src", "lev_src_inc", "lev_src_dec"):
            o[k] = np.zeros((ncol, nlay, ngpt), order="F")
        o["sfc_src"] = np.zeros((ncol, ngpt), order="F")
        o["sfc_src_jac"] = np.zeros((ncol, ngpt), order="F")
    else:
        o["ssa"] = np.zeros((ncol, nlay, ngpt), order="F")
        o["g"] = np.zeros((ncol, nlay, ngpt), order="F")
        o["toa_src"] = np.zeros((ncol, ngpt), order="F")
    if interp:
        nf = kd.nflav
        o["jtemp"] = np.zeros((ncol, nlay), dtype=np.int32, order="F")
        o["jpress"] = np.zeros((ncol, nlay), dtype=np.int32, order="F")
        o["tropo"] = np.zeros((ncol, nlay), dtype=np.int8, order="F")
        o["jeta"] = np.zeros((2, ncol, nlay, nf), dtype=np.int32, order="F")
        o["col_mix"] = np.zeros((2, ncol, nlay, nf), order="F")
        o["fmajor"] = np.zeros((2, 2, 2, ncol, nlay, nf), order="F")
        o["fminor"] = np.zeros((2, 2, ncol, nlay, nf), order="F")
    t_sfc = _f64(atm.t_sfc)
    t_lev = _f64(atm.t_lev)
    rc = lib().oracle_gas_optics(
        kd.h, ncol, nlay, ptr(_f64(atm.p_lay)), ptr(_f64(atm.p_lev)), ptr(_f64(atm.t_lay)),
        ptr(t_sfc), ptr(t_lev), C.byref(gc.s), ptr(col_dry), ptr(o["tau"]), ptr(o.get("ssa")),
        ptr(o.get("g")), ptr(o.get("lay_src")), ptr(o.get("lev_src_inc")),
        ptr(o.get("lev_src_dec")), ptr(o.get("sfc_src")), ptr(o.get("sfc_src_jac")),
        ptr(o.get("toa_src")), ptr(o.get("jtemp"), C.c_int), ptr(o.get("jpress"), C.c_int),
        ptr(o.get("tropo"), C.c_int8), ptr(o.get("jeta"), C.c_int), ptr(o.get("col_mix")),
        ptr(o.get("fmajor")), ptr(o.get("fminor")))
    _check(rc)
    return o


class Cloud:
    def __init__(self, lut):
        self.lut = lut
        self.nbnd = lut.nbnd
        self._keep = [_f64(a) for a in (lut.lut_extliq, lut.lut_ssaliq, lut.lut_asyliq,
                                        lut.lut_extice, lut.lut_ssaice, lut.lut_asyice)]
        k = self._keep
        h = lib().oracle_cloud_create(
            lut.nbnd, lut.radliq_lwr, lut.radliq_upr, lut.radice_lwr, lut.radice_upr,
            lut.lut_extliq.shape[0], ptr(k[0]), ptr(k[1]), ptr(k[2]),
            lut.lut_extice.shape[0], ptr(k[3]), ptr(k[4]), ptr(k[5]))
        self.h = C.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_cloud_destroy(self.h)
            self.h = None

    def optics(self, lwp, iwp, rel, rei, two_stream):
        lwp, iwp, rel, rei = (_f64(a) for a in (lwp, iwp, rel, rei))
        ncol, nlay = lwp.shape
        tau = np.zeros((ncol, nlay, self.nbnd), order="F")
        ssa = np.zeros_like(tau) if two_stream else None
        g = np.zeros_like(tau) if two_stream else None
        _check(lib().oracle_cloud_optics(self.h, ncol, nlay, ptr(lwp), ptr(iwp), ptr(rel), ptr(rei),
                                         int(two_stream), ptr(tau), ptr(ssa), ptr(g)))
        return (tau, ssa, g) if two_stream else tau


AEROSOL_FIELDS = tuple(f"aermr{i:02d}" for i in range(1, 12))


def aerosol_optics(aermr, rh, plev, lut):
    """Aerosol_optics::aerosol_optics: aermr maps aermr01..aermr11 to (ncol, nlay) mass mixing ratios;
    lut holds rh_upper (nhum), mext/ssa/g_phobic (nbnd, nphobic), mext/ssa/g_philic (nbnd, nhum, nphilic).
    Returns tau, ssa, g (ncol, nlay, nbnd)."""
    rh, plev = _f64(rh), _f64(plev)
    ncol, nlay = rh.shape
    fields = np.stack([np.broadcast_to(np.asarray(aermr[n], dtype=np.float64), (ncol, nlay)).T
                       for n in AEROSOL_FIELDS])           # (11, nlay, ncol) C order = 11 Fortran fields
    fields = np.ascontiguousarray(fields)
    tabs = [_f64(lut[n]) for n in ("rh_upper", "mext_phobic", "ssa_phobic", "g_phobic",
                                   "mext_philic", "ssa_philic", "g_philic")]
    nbnd, nphobic = tabs[1].shape
    _, nhum, nphilic = tabs[4].shape
    tau, ssa, g = (np.zeros((ncol, nlay, nbnd), order="F") for _ in range(3))
    _check(lib().oracle_aerosol_optics(ncol, nlay, nbnd, nhum, nphobic, nphilic, ptr(fields), ptr(rh), ptr(plev),
                                       *[ptr(t) for t in tabs], ptr(tau), ptr(ssa), ptr(g)))
    return tau, ssa, g


def delta_scale(tau, ssa, g):
    tau, ssa, g = (np.array(a, dtype=np.float64, order="F") for a in (tau, ssa, g))
    ncol, nlay, ngpt = tau.shape
    lib().oracle_delta_scale_2str(ncol, nlay, ngpt, ptr(tau), ptr(ssa), ptr(g))
    return tau, ssa, g


def inc_bybnd(tau1, tau2, band_lims, ssa1=None, g1=None, ssa2=None, g2=None):
    tau1 = np.array(tau1, dtype=np.float64, order="F")
    ssa1 = None if ssa1 is None else np.array(ssa1, dtype=np.float64, order="F")
    g1 = None if g1 is None else np.array(g1, dtype=np.float64, order="F")
    tau2 = _f64(tau2)
    ssa2 = None if ssa2 is None else _f64(ssa2)
    g2 = None if g2 is None else _f64(g2)
    band_lims = _i32(band_lims)
    ncol, nlay, ngpt = tau1.shape
    lib().oracle_inc_bybnd(ncol, nlay, ngpt, ptr(tau1), ptr(ssa1), ptr(g1), ptr(tau2), ptr(ssa2),
                           ptr(g2), band_lims.shape[1], ptr(band_lims, C.c_int))
    return (tau1, ssa1, g1) if ssa1 is not None else tau1


def increment(tau1, tau2, ssa1=None, g1=None, ssa2=None, g2=None):
    tau1 = np.array(tau1, dtype=np.float64, order="F")
    ssa1 = None if ssa1 is None else np.array(ssa1, dtype=np.float64, order="F")
    g1 = None if g1 is None else np.array(g1, dtype=np.float64, order="F")
    ncol, nlay, ngpt = tau1.shape
    lib().oracle_increment(ncol, nlay, ngpt, ptr(tau1), ptr(ssa1), ptr(g1), ptr(_f64(tau2)),
                           ptr(None if ssa2 is None else _f64(ssa2)),
                           ptr(None if g2 is None else _f64(g2)))
    return (tau1, ssa1, g1) if ssa1 is not None else tau1


def rte_lw(tau, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis, band_lims, top_at_1,
           inc_flux=None, n_gauss_angles=1, do_broadband=False, sfc_src_jac=None):
    tau, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis = (
        _f64(a) for a in (tau, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis))
    band_lims = _i32(band_lims)
    ncol, nlay, ngpt = tau.shape
    nlev = nlay + 1
    shape = (ncol, nlev) if do_broadband else (ncol, nlev, ngpt)
    up = np.zeros(shape, order="F")
    dn = np.zeros(shape, order="F")
    jac = np.zeros((ncol, nlev), order="F") if sfc_src_jac is not None else None
    _check(lib().oracle_rte_lw(ncol, nlay, ngpt, band_lims.shape[1], ptr(band_lims, C.c_int),
                               int(top_at_1), ptr(tau), ptr(lay_src), ptr(lev_inc), ptr(lev_dec),
                               ptr(sfc_src), ptr(sfc_emis), ptr(None if inc_flux is None else _f64(inc_flux)),
                               n_gauss_angles, int(do_broadband), ptr(up), ptr(dn),
                               ptr(None if sfc_src_jac is None else _f64(sfc_src_jac)), ptr(jac)))
    return (up, dn, jac) if jac is not None else (up, dn)


def rte_lw_rescaled(tau, ssa, g, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis, band_lims, top_at_1,
                    inc_flux=None, n_gauss_angles=1, do_broadband=False, sfc_src_jac=None):
    """rte_lw with the Tang et al. (2018) scattering rescaling (do_rescaling = true)."""
    tau, ssa, g, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis = (
        _f64(a) for a in (tau, ssa, g, lay_src, lev_inc, lev_dec, sfc_src, sfc_emis))
    band_lims = _i32(band_lims)
    ncol, nlay, ngpt = tau.shape
    nlev = nlay + 1
    shape = (ncol, nlev) if do_broadband else (ncol, nlev, ngpt)
    up = np.zeros(shape, order="F")
    dn = np.zeros(shape, order="F")
    jac = np.zeros((ncol, nlev), order="F") if sfc_src_jac is not None else None
    _check(lib().oracle_rte_lw_rescaled(ncol, nlay, ngpt, band_lims.shape[1], ptr(band_lims, C.c_int),
                                        int(top_at_1), ptr(tau), ptr(ssa), ptr(g), ptr(lay_src), ptr(lev_inc),
                                        ptr(lev_dec), ptr(sfc_src), ptr(sfc_emis),
                                        ptr(None if inc_flux is None else _f64(inc_flux)), n_gauss_angles,
                                        int(do_broadband), ptr(up), ptr(dn),
                                        ptr(None if sfc_src_jac is None else _f64(sfc_src_jac)), ptr(jac)))
    return (up, dn, jac) if jac is not None else (up, dn)


def lw_solver_2stream(tau, ssa, g, lev_inc, lev_dec, sfc_emis, sfc_src, top_at_1, inc_flux=None):
    """lw_solver_2stream: spectral fluxes (ncol, nlay+1, ngpt); sfc_emis, sfc_src, inc_flux (ncol, ngpt)."""
    tau, ssa, g, lev_inc, lev_dec, sfc_emis, sfc_src = (
        _f64(a) for a in (tau, ssa, g, lev_inc, lev_dec, sfc_emis, sfc_src))
    ncol, nlay, ngpt = tau.shape
    up = np.zeros((ncol, nlay + 1, ngpt), order="F")
    dn = np.zeros((ncol, nlay + 1, ngpt), order="F")
    _check(lib().oracle_lw_solver_2stream(ncol, nlay, ngpt, int(top_at_1), ptr(tau), ptr(ssa), ptr(g),
                                          ptr(lev_inc), ptr(lev_dec), ptr(sfc_emis), ptr(sfc_src),
                                          ptr(None if inc_flux is None else _f64(inc_flux)), ptr(up), ptr(dn)))
    return up, dn


def rte_sw(tau, ssa, g, mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif, band_lims, top_at_1,
           inc_flux_dif=None, do_broadband=False):
    tau, ssa, g, mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif = (
        _f64(a) for a in (tau, ssa, g, mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif))
    band_lims = _i32(band_lims)
    ncol, nlay, ngpt = tau.shape
    nlev = nlay + 1
    shape = (ncol, nlev) if do_broadband else (ncol, nlev, ngpt)
    up, dn, dr = (np.zeros(shape, order="F") for _ in range(3))
    _check(lib().oracle_rte_sw(ncol, nlay, ngpt, band_lims.shape[1], ptr(band_lims, C.c_int),
                               int(top_at_1), ptr(tau), ptr(ssa), ptr(g), ptr(mu0), ptr(inc_flux_dir),
                               ptr(sfc_alb_dir), ptr(sfc_alb_dif),
                               ptr(None if inc_flux_dif is None else _f64(inc_flux_dif)),
                               int(do_broadband), ptr(up), ptr(dn), ptr(dr)))
    return up, dn, dr


def sw_two_stream(tau, w0, g, mu0):
    tau, w0, g, mu0 = (np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()) for a in (tau, w0, g, mu0))
    n = tau.size
    out = [np.zeros(n) for _ in range(5)]
    lib().oracle_sw_two_stream(n, ptr(tau), ptr(w0), ptr(g), ptr(mu0), *[ptr(o) for o in out])
    return out


def sum_broadband(gpt_flux):
    gpt_flux = _f64(gpt_flux)
    ncol, nlev, ngpt = gpt_flux.shape
    out = np.zeros((ncol, nlev), order="F")
    lib().oracle_sum_broadband(ncol, nlev, ngpt, ptr(gpt_flux), ptr(out))
    return out


def sum_byband(gpt_flux, band_lims):
    gpt_flux, band_lims = _f64(gpt_flux), _i32(band_lims)
    ncol, nlev, ngpt = gpt_flux.shape
    nbnd = band_lims.shape[1]
    out = np.zeros((ncol, nlev, nbnd), order="F")
    lib().oracle_sum_byband(ncol, nlev, nbnd, ptr(band_lims, C.c_int), ptr(gpt_flux), ptr(out))
    return out


def heating_rate(flux_net, plev):
    flux_net, plev = _f64(flux_net), _f64(plev)
    ncol, nlev = flux_net.shape
    out = np.zeros((ncol, nlev - 1), order="F")
    lib().oracle_heating_rate(ncol, nlev - 1, ptr(flux_net), ptr(plev), ptr(out))
    return out


def set_threads(n: int) -> int:
    """Threads of solve_lw/solve_sw's column-block loop (1 = serial, as the reference driver);
    returns the previous value."""
    return int(lib().oracle_set_threads(int(n)))


def solve_lw(kd: Kdist, atm, cloud: Cloud | None = None, bnd=False, col_dry=None, n_col_block=12):
    """Radiation_solver_longwave::solve restated; returns dict of (ncol, nlev[, nbnd]) arrays."""
    ncol, nlay = atm.p_lay.shape
    nlev = nlay + 1
    gc = GasConcs(atm.vmr)
    o = {k: np.zeros((ncol, nlev), order="F") for k in ("flux_up", "flux_dn", "flux_net")}
    if bnd:
        for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_net"):
            o[k] = np.zeros((ncol, nlev, kd.nbnd), order="F")
    cl = [_f64(getattr(atm, k)) if cloud is not None else None for k in ("lwp", "iwp", "rel", "rei")]
    _check(lib().oracle_solve_lw(
        kd.h, cloud.h if cloud is not None else None, ncol, nlay, ptr(_f64(atm.p_lay)),
        ptr(_f64(atm.p_lev)), ptr(_f64(atm.t_lay)), ptr(_f64(atm.t_lev)), ptr(_f64(atm.t_sfc)),
        C.byref(gc.s), ptr(None if col_dry is None else _f64(col_dry)), ptr(_f64(atm.emis_sfc)),
        *[ptr(a) for a in cl], ptr(o["flux_up"]), ptr(o["flux_dn"]), ptr(o["flux_net"]),
        ptr(o.get("bnd_flux_up")), ptr(o.get("bnd_flux_dn")), ptr(o.get("bnd_flux_net")),
        n_col_block))
    return o


def solve_sw(kd: Kdist, atm, cloud: Cloud | None = None, delta_cloud=True, bnd=False,
             col_dry=None, n_col_block=12, tsi_scaling=None):
    ncol, nlay = atm.p_lay.shape
    nlev = nlay + 1
    gc = GasConcs(atm.vmr)
    if tsi_scaling is None:
        tsi_scaling = np.full(ncol, atm.tsi / kd.scalar("tsi"))
    o = {k: np.zeros((ncol, nlev), order="F") for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net")}
    if bnd:
        for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_dn_dir", "bnd_flux_net"):
            o[k] = np.zeros((ncol, nlev, kd.nbnd), order="F")
    cl = [_f64(getattr(atm, k)) if cloud is not None else None for k in ("lwp", "iwp", "rel", "rei")]
    _check(lib().oracle_solve_sw(
        kd.h, cloud.h if cloud is not None else None, int(delta_cloud), ncol, nlay,
        ptr(_f64(atm.p_lay)), ptr(_f64(atm.p_lev)), ptr(_f64(atm.t_lay)), C.byref(gc.s),
        ptr(None if col_dry is None else _f64(col_dry)), ptr(_f64(atm.sfc_alb)),
        ptr(_f64(atm.sfc_alb)), ptr(_f64(tsi_scaling)), ptr(_f64(atm.mu0)),
        *[ptr(a) for a in cl], ptr(o["flux_up"]), ptr(o["flux_dn"]), ptr(o["flux_dn_dir"]),
        ptr(o["flux_net"]), ptr(o.get("bnd_flux_up")), ptr(o.get("bnd_flux_dn")),
        ptr(o.get("bnd_flux_dn_dir")), ptr(o.get("bnd_flux_net")), n_col_block))
    return o
