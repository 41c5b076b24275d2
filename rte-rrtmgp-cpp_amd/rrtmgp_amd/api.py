"""Python binding of the class API through the C facade (include/rrtmgp_api.h).

Objects mirror the reference classes: KDist ~ Gas_optics_rrtmgp_gpu, Cloud ~ Cloud_optics_gpu,
GasConcs ~ Gas_concs_gpu, SolverLW/SolverSW ~ Radiation_solver_{longwave,shortwave}_gpu; the
free functions wrap Rte_lw_gpu::rte_lw, Rte_sw_gpu::rte_sw, add_to, delta_scale and the flux
reductions.  Arrays are DeviceArrays (or numpy arrays, uploaded on the fly) in Fortran layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_int8)


class KdistFile(C.Structure):
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


class CloudLutC(C.Structure):
    _fields_ = [("nbnd", C.c_int), ("band_lims_wvn", _dp),
                ("radliq_lwr", C.c_double), ("radliq_upr", C.c_double), ("radliq_fac", C.c_double),
                ("radice_lwr", C.c_double), ("radice_upr", C.c_double), ("radice_fac", C.c_double),
                ("nsize_liq", C.c_int), ("lut_extliq", _dp), ("lut_ssaliq", _dp), ("lut_asyliq", _dp),
                ("nsize_ice", C.c_int), ("lut_extice", _dp), ("lut_ssaice", _dp), ("lut_asyice", _dp)]


class AerosolLutC(C.Structure):
    _fields_ = [("nbnd", C.c_int), ("nhum", C.c_int), ("nphobic", C.c_int), ("nphilic", C.c_int),
                ("band_lims_wvn", _dp), ("rh_upper", _dp), ("mext_phobic", _dp), ("ssa_phobic", _dp),
                ("g_phobic", _dp), ("mext_philic", _dp), ("ssa_philic", _dp), ("g_philic", _dp)]


class ApiGases(C.Structure):
    _fields_ = [("n", C.c_int), ("names", C.c_char_p), ("data", C.POINTER(C.c_void_p)),
                ("dims", C.POINTER(C.c_int64))]


def _names(lst):
    return "\n".join(lst).encode()


def _ptr(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def real_dtype(precision):
    return np.float64 if precision == 8 else np.float32


class _Handle:
    precision = 8

    def _check(self, rc):
        _lib.check_api(rc, self.precision)

    @property
    def L(self):
        return _lib.cpp(self.precision)


def _dev(x, dtype):
    """DeviceArray of dtype from a DeviceArray (checked) or a host array (uploaded)."""
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        if x.dtype != np.dtype(dtype):
            raise TypeError(f"device array has dtype {x.dtype}, expected {np.dtype(dtype)}")
        return x
    return DeviceArray.from_host(np.asarray(x), dtype=dtype)


def _p(d):
    return None if d is None else d.ptr


class KDist(_Handle):
    """Gas_optics_rrtmgp_gpu built from a synthetic.KDist (coefficient-file layout)."""

    def __init__(self, kd, available_gases, precision=8):
        self.precision = precision
        self.src = kd
        keep = []

        def hold(a, dtype):
            a = np.asfortranarray(np.asarray(a, dtype=dtype))
            keep.append(a)
            return a

        f = KdistFile()
        f.ngas = len(kd.gas_names)
        f.gas_names = _names(kd.gas_names)
        f.nbnd, f.ngpt = kd.nbnd, kd.ngpt
        f.ntemp, f.npres, f.neta = kd.temp_ref.size, kd.press_ref.size, kd.kmajor.shape[1]
        f.nplanck = kd.totplnk.shape[0] if kd.totplnk is not None else 0
        f.key_species = _ptr(hold(kd.key_species, np.int32), C.c_int)
        f.band2gpt = _ptr(hold(kd.band2gpt, np.int32), C.c_int)
        f.band_lims_wavenum = _ptr(hold(kd.band_lims_wavenum, np.float64))
        f.press_ref = _ptr(hold(kd.press_ref, np.float64))
        f.press_ref_trop = kd.press_ref_trop
        f.temp_ref = _ptr(hold(kd.temp_ref, np.float64))
        f.temp_ref_p, f.temp_ref_t = kd.temp_ref_p, kd.temp_ref_t
        f.vmr_ref = _ptr(hold(kd.vmr_ref, np.float64))
        f.kmajor = _ptr(hold(kd.kmajor, np.float64))
        f.nminor_ident = len(kd.identifier_minor)
        f.gas_minor = _names(kd.gas_minor)
        f.identifier_minor = _names(kd.identifier_minor)
        for side in ("lower", "upper"):
            km = getattr(kd, f"kminor_{side}")
            setattr(f, f"nminor_{side}", len(getattr(kd, f"minor_gases_{side}")))
            setattr(f, f"ncontrib_{side}", km.shape[0])
            setattr(f, f"kminor_{side}", _ptr(hold(km, np.float64)))
            setattr(f, f"minor_gases_{side}", _names(getattr(kd, f"minor_gases_{side}")))
            setattr(f, f"minor_limits_gpt_{side}", _ptr(hold(getattr(kd, f"minor_limits_gpt_{side}"), np.int32), C.c_int))
            setattr(f, f"minor_scales_with_density_{side}", _ptr(hold(getattr(kd, f"minor_scales_with_density_{side}"), np.int8), C.c_int8))
            setattr(f, f"scaling_gas_{side}", _names(getattr(kd, f"scaling_gas_{side}")))
            setattr(f, f"scale_by_complement_{side}", _ptr(hold(getattr(kd, f"scale_by_complement_{side}"), np.int8), C.c_int8))
            setattr(f, f"kminor_start_{side}", _ptr(hold(getattr(kd, f"kminor_start_{side}"), np.int32), C.c_int))
        if kd.totplnk is not None:
            f.totplnk = _ptr(hold(kd.totplnk, np.float64))
            f.planck_frac = _ptr(hold(kd.planck_frac, np.float64))
        if kd.solar_src_quiet is not None:
            f.solar_src_quiet = _ptr(hold(kd.solar_src_quiet, np.float64))
            f.solar_src_facular = _ptr(hold(kd.solar_src_facular, np.float64))
            f.solar_src_sunspot = _ptr(hold(kd.solar_src_sunspot, np.float64))
            f.tsi_default, f.mg_default, f.sb_default = kd.tsi_default, kd.mg_default, kd.sb_default
            f.rayl_lower = _ptr(hold(kd.rayl_lower, np.float64))
            f.rayl_upper = _ptr(hold(kd.rayl_upper, np.float64))
        f.n_available = len(available_gases)
        f.available_gases = _names(list(available_gases))
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_kdist_create(C.byref(f), C.byref(h)))
        self.h = h
        info = np.zeros(12, dtype=np.int64)
        self._check(self.L.rrtmgp_api_kdist_info(self.h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        (self.ngas, self.nflav, self.neta, self.npres, self.ntemp, self.ngpt, self.nbnd, self.nplanck,
         self.nminor_lower, self.nminor_upper, self.idx_h2o, is_lw) = [int(x) for x in info]
        self.is_lw = bool(is_lw)
        buf = C.create_string_buffer(4096)
        self._check(self.L.rrtmgp_api_kdist_gas_names(self.h, buf, 4096))
        self.gas_names = buf.value.decode().split("\n")
        self.band_lims = np.asfortranarray(kd.band2gpt.astype(np.int32))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rrtmgp_api_kdist_destroy(self.h)
            self.h = None

    def get_int(self, name, cap=1 << 20):
        out = np.zeros(cap, dtype=np.int32)
        n = C.c_int64(0)
        self._check(self.L.rrtmgp_api_kdist_get_int(self.h, name.encode(), _ptr(out, C.c_int),
                                                    C.c_int64(cap), C.byref(n)))
        return out[:n.value].copy()

    def get_real(self, name, cap=1 << 22):
        out = np.zeros(cap)
        n = C.c_int64(0)
        self._check(self.L.rrtmgp_api_kdist_get_real(self.h, name.encode(), _ptr(out), C.c_int64(cap), C.byref(n)))
        return out[:n.value].copy()

    def scalar(self, name):
        return float(self.get_real(name, 1)[0])

    def set_check_ranges(self, on):
        self._check(self.L.rrtmgp_api_kdist_set_check_ranges(self.h, int(on)))


class GasConcs(_Handle):
    """Gas_concs_gpu from a name -> vmr dict ((ncol, nlay) arrays or (1, 1) scalars)."""

    def __init__(self, vmr: dict, precision=8):
        self.precision = precision
        dt = real_dtype(precision)
        self.names = list(vmr.keys())
        self.arrays = [_dev(vmr[k] if isinstance(vmr[k], DeviceArray) else np.atleast_2d(vmr[k]), dt)
                       for k in self.names]
        dims = np.array([d for a in self.arrays for d in a.shape], dtype=np.int64)
        ptrs = (C.c_void_p * len(self.arrays))(*[a.ptr.value for a in self.arrays])
        g = ApiGases(len(self.names), _names(self.names), ptrs, dims.ctypes.data_as(C.POINTER(C.c_int64)))
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_gas_concs_create(C.byref(g), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rrtmgp_api_gas_concs_destroy(self.h)
            self.h = None


class Cloud(_Handle):
    """Cloud_optics_gpu from a synthetic.CloudLUT."""

    def __init__(self, lut, precision=8):
        self.precision = precision
        self._keep = [np.asfortranarray(np.asarray(a, dtype=np.float64)) for a in (
            lut.band_lims_wvn, lut.lut_extliq, lut.lut_ssaliq, lut.lut_asyliq, lut.lut_extice,
            lut.lut_ssaice, lut.lut_asyice)]
        k = self._keep
        s = CloudLutC(lut.nbnd, _ptr(k[0]), lut.radliq_lwr, lut.radliq_upr, lut.radliq_fac,
                      lut.radice_lwr, lut.radice_upr, lut.radice_fac, lut.lut_extliq.shape[0],
                      _ptr(k[1]), _ptr(k[2]), _ptr(k[3]), lut.lut_extice.shape[0], _ptr(k[4]),
                      _ptr(k[5]), _ptr(k[6]))
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_cloud_create(C.byref(s), C.byref(h)))
        self.h = h
        self.nbnd = lut.nbnd

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rrtmgp_api_cloud_destroy(self.h)
            self.h = None

    def optics(self, lwp, iwp, rel, rei, two_stream, delta_scale=False, coalbedo=False):
        """tau (and with two_stream ssa, g) per band; with coalbedo the second array is 1 - ssa."""
        dt = real_dtype(self.precision)
        lwp, iwp, rel, rei = (_dev(a, dt) for a in (lwp, iwp, rel, rei))
        ncol, nlay = lwp.shape
        tau = DeviceArray((ncol, nlay, self.nbnd), dt)
        ssa = DeviceArray((ncol, nlay, self.nbnd), dt) if two_stream else None
        g = DeviceArray((ncol, nlay, self.nbnd), dt) if two_stream else None
        self._check(self.L.rrtmgp_api_cloud_optics(self.h, ncol, nlay, lwp.ptr, iwp.ptr, rel.ptr, rei.ptr,
                                                   2 if two_stream and coalbedo else int(two_stream),
                                                   int(delta_scale), tau.ptr, _p(ssa), _p(g)))
        return (tau, ssa, g) if two_stream else tau


AEROSOL_FIELDS = tuple(f"aermr{i:02d}" for i in range(1, 12))


class Aerosol(_Handle):
    """Aerosol_optics_gpu from a dict of the tables load_and_init_aerosol_optics reads
    (src_test/Radiation_solver.cpp:329-366): rh_upper (nhum), mext/ssa/g_phobic (nbnd, nphobic),
    mext/ssa/g_philic (nbnd, nhum, nphilic); band_lims_wvn (2, nbnd) is optional (the reference's
    loader leaves it unset)."""

    def __init__(self, lut: dict, precision=8):
        self.precision = precision
        names = ("rh_upper", "mext_phobic", "ssa_phobic", "g_phobic", "mext_philic", "ssa_philic", "g_philic")
        k = {n: np.asfortranarray(np.asarray(lut[n], dtype=np.float64)) for n in names}
        nbnd, nphobic = k["mext_phobic"].shape
        _, nhum, nphilic = k["mext_philic"].shape
        blw = lut.get("band_lims_wvn")
        k["band_lims_wvn"] = np.asfortranarray(np.asarray(
            blw if blw is not None else np.stack([np.arange(nbnd), np.arange(1, nbnd + 1)]), dtype=np.float64))
        self._keep = k
        s = AerosolLutC(nbnd, nhum, nphobic, nphilic, _ptr(k["band_lims_wvn"]), *[_ptr(k[n]) for n in names])
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_aerosol_create(C.byref(s), C.byref(h)))
        self.h = h
        self.nbnd = nbnd

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rrtmgp_api_aerosol_destroy(self.h)
            self.h = None

    def optics(self, concs: GasConcs, rh, plev, delta_scale=False, coalbedo=False):
        """tau, ssa, g (ncol, nlay, nbnd) of the aerosol mixing ratios held by concs; with coalbedo
        the second array is 1 - ssa."""
        dt = real_dtype(self.precision)
        rh, plev = _dev(rh, dt), _dev(plev, dt)
        ncol, nlay = rh.shape
        out = [DeviceArray((ncol, nlay, self.nbnd), dt) for _ in range(3)]
        self._check(self.L.rrtmgp_api_aerosol_optics(self.h, ncol, nlay, concs.h, rh.ptr, plev.ptr,
                                                     int(delta_scale), int(coalbedo), *[o.ptr for o in out]))
        return tuple(out)


def get_col_dry(vmr_h2o, plev, precision=8):
    dt = real_dtype(precision)
    h2o, plev = _dev(vmr_h2o, dt), _dev(plev, dt)
    ncol, nlev = plev.shape
    out = DeviceArray((ncol, nlev - 1), dt)
    _lib.check_api(_lib.cpp(precision).rrtmgp_api_get_col_dry(ncol, nlev - 1, h2o.ptr, C.c_int64(h2o.shape[0]),
                                                              C.c_int64(h2o.shape[1]), plev.ptr, out.ptr), precision)
    return out


def gas_optics(kd: KDist, atm, gas_concs: GasConcs, col_dry, coalbedo=False):
    """Gas_optics_rrtmgp_gpu::gas_optics (LW or SW by k-distribution); returns DeviceArrays.  With
    coalbedo (SW) o["coalbedo"] = 1 - ssa replaces o["ssa"]."""
    dt = real_dtype(kd.precision)
    ncol, nlay = atm.p_lay.shape
    ng = kd.ngpt
    play, plev, tlay = (_dev(a, dt) for a in (atm.p_lay, atm.p_lev, atm.t_lay))
    cd = _dev(col_dry, dt)
    o = {"tau": DeviceArray((ncol, nlay, ng), dt)}
    if kd.is_lw:
        tlev, tsfc = _dev(atm.t_lev, dt), _dev(atm.t_sfc, dt)
        for k in ("lay_src", "lev_src_inc", "lev_src_dec"):
            o[k] = DeviceArray((ncol, nlay, ng), dt)
        o["sfc_src"] = DeviceArray((ncol, ng), dt)
        o["sfc_src_jac"] = DeviceArray((ncol, ng), dt)
        kd._check(kd.L.rrtmgp_api_gas_optics_lw(
            kd.h, ncol, nlay, play.ptr, plev.ptr, tlay.ptr, tsfc.ptr, tlev.ptr, gas_concs.h, cd.ptr,
            o["tau"].ptr, o["lay_src"].ptr, o["lev_src_inc"].ptr, o["lev_src_dec"].ptr,
            o["sfc_src"].ptr, o["sfc_src_jac"].ptr))
    else:
        key = "coalbedo" if coalbedo else "ssa"
        o[key] = DeviceArray((ncol, nlay, ng), dt)
        o["g"] = DeviceArray((ncol, nlay, ng), dt)
        o["toa_src"] = DeviceArray((ncol, ng), dt)
        fn = kd.L.rrtmgp_api_gas_optics_sw_coalbedo if coalbedo else kd.L.rrtmgp_api_gas_optics_sw
        kd._check(fn(kd.h, ncol, nlay, play.ptr, plev.ptr, tlay.ptr, gas_concs.h, cd.ptr, o["tau"].ptr,
                     o[key].ptr, o["g"].ptr, o["toa_src"].ptr))
    return o


def add_to(kd: KDist, tau, tau_in, ssa=None, g=None, ssa_in=None, g_in=None, coalbedo=False):
    """add_to(op_inout, op_in): same spectral grid or band-resolved input; in place.  With coalbedo
    ssa and ssa_in hold 1 - ssa and the input is band-resolved."""
    ncol, nlay, _ = tau.shape
    if coalbedo:
        kd._check(kd.L.rrtmgp_api_add_to_coalbedo(kd.h, ncol, nlay, tau.ptr, ssa.ptr, g.ptr, tau_in.ptr,
                                                  ssa_in.ptr, g_in.ptr))
        return
    kd._check(kd.L.rrtmgp_api_add_to(kd.h, ncol, nlay, tau_in.shape[2], tau.ptr, _p(ssa), _p(g),
                                     tau_in.ptr, _p(ssa_in), _p(g_in)))


def delta_scale(tau, ssa, g, precision=8):
    ncol, nlay, ngpt = tau.shape
    _lib.check_api(_lib.cpp(precision).rrtmgp_api_delta_scale(ncol, nlay, ngpt, tau.ptr, ssa.ptr, g.ptr), precision)


def rte_lw(kd: KDist, top_at_1, tau, lay_src, lev_src_inc, lev_src_dec, sfc_src, sfc_emis,
           inc_flux=None, n_gauss_angles=1, do_broadband=False):
    dt = real_dtype(kd.precision)
    ncol, nlay, ng = tau.shape
    emis = _dev(sfc_emis, dt)
    inc = _dev(inc_flux, dt)
    shape = (ncol, nlay + 1) if do_broadband else (ncol, nlay + 1, ng)
    up, dn = DeviceArray(shape, dt), DeviceArray(shape, dt)
    kd._check(kd.L.rrtmgp_api_rte_lw(kd.h, ncol, nlay, int(top_at_1), tau.ptr, lay_src.ptr, lev_src_inc.ptr,
                                     lev_src_dec.ptr, sfc_src.ptr, emis.ptr, _p(inc), n_gauss_angles,
                                     int(do_broadband), up.ptr, dn.ptr))
    return up, dn


def rte_sw(kd: KDist, top_at_1, tau, ssa, g, mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif,
           inc_flux_dif=None, do_broadband=False):
    dt = real_dtype(kd.precision)
    ncol, nlay, ng = tau.shape
    mu0, inc, ad, af = (_dev(a, dt) for a in (mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif))
    dif = _dev(inc_flux_dif, dt)
    shape = (ncol, nlay + 1) if do_broadband else (ncol, nlay + 1, ng)
    up, dn, dr = DeviceArray(shape, dt), DeviceArray(shape, dt), DeviceArray(shape, dt)
    kd._check(kd.L.rrtmgp_api_rte_sw(kd.h, ncol, nlay, int(top_at_1), tau.ptr, ssa.ptr, g.ptr, mu0.ptr,
                                     inc.ptr, ad.ptr, af.ptr, _p(dif), int(do_broadband), up.ptr, dn.ptr, dr.ptr))
    return up, dn, dr


def rte_sw_bnd(kd: KDist, top_at_1, tau, ssa, g, mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif, p_lev=None,
               coalbedo=False):
    """Broadband SW fluxes summed inside the solver (rrtmgp_hip_sw_solver_2stream_bnd_*), with
    albedos (nbnd, ncol); given p_lev (ncol, nlev), also heating rates (ncol, nlay) [K/day] from the
    layers' absorbed flux summed over g-points in FP64.  With coalbedo ssa holds 1 - ssa.  Returns a
    dict of DeviceArrays."""
    dt = real_dtype(kd.precision)
    ncol, nlay, ng = tau.shape
    mu0, inc, ad, af, pl = (_dev(a, dt) for a in (mu0, inc_flux_dir, sfc_alb_dir, sfc_alb_dif, p_lev))
    o = {k: DeviceArray((ncol, nlay + 1), dt) for k in ("flux_up", "flux_dn", "flux_dn_dir")}
    o["heating_rate"] = DeviceArray((ncol, nlay), dt) if pl is not None else None
    fn = getattr(_lib.hip(), "rrtmgp_hip_sw_solver_2stream_bnd_" + ("coalbedo_" if coalbedo else "")
                 + ("f64" if kd.precision == 8 else "f32"))
    bl = kd.band_lims
    i = C.c_int64
    _lib.check_hip(fn(i(ncol), i(nlay), i(ng), C.c_int(int(top_at_1)), tau.ptr, ssa.ptr, g.ptr, mu0.ptr, ad.ptr,
                      af.ptr, inc.ptr, C.c_int(0), None, i(kd.nbnd), bl.ctypes.data_as(C.POINTER(C.c_int)), None,
                      None, None, o["flux_up"].ptr, o["flux_dn"].ptr, o["flux_dn_dir"].ptr, _p(pl),
                      _p(o["heating_rate"]), None))
    return {k: v for k, v in o.items() if v is not None}


def fluxes_reduce(kd: KDist, gpt_up, gpt_dn, gpt_dir=None, byband=False):
    dt = real_dtype(kd.precision)
    ncol, nlev, _ = gpt_up.shape
    o = {k: DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_net")}
    o["flux_dn_dir"] = DeviceArray((ncol, nlev), dt) if gpt_dir is not None else None
    for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_net", "bnd_flux_dn_dir"):
        o[k] = DeviceArray((ncol, nlev, kd.nbnd), dt) if byband else None
    if gpt_dir is None:
        o["bnd_flux_dn_dir"] = None
    kd._check(kd.L.rrtmgp_api_fluxes_reduce(
        kd.h, ncol, nlev, gpt_up.ptr, gpt_dn.ptr, _p(gpt_dir), o["flux_up"].ptr, o["flux_dn"].ptr,
        _p(o["flux_dn_dir"]), o["flux_net"].ptr, _p(o["bnd_flux_up"]), _p(o["bnd_flux_dn"]),
        _p(o["bnd_flux_dn_dir"]), _p(o["bnd_flux_net"])))
    return {k: v for k, v in o.items() if v is not None}


def heating_rate(flux_up, flux_dn, plev, precision=8):
    dt = real_dtype(precision)
    up, dn, p = (_dev(a, dt) for a in (flux_up, flux_dn, plev))
    ncol, nlev = up.shape
    out = DeviceArray((ncol, nlev - 1), dt)
    _lib.check_api(_lib.cpp(precision).rrtmgp_api_heating_rate(ncol, nlev - 1, up.ptr, dn.ptr, p.ptr, out.ptr),
                   precision)
    return out


class _Solver(_Handle):
    def __del__(self):
        if getattr(self, "h", None):
            self.L.rrtmgp_api_solver_destroy(self.h)
            self.h = None

    def set_block(self, n):
        self._check(self.L.rrtmgp_api_solver_set_block(self.h, int(n)))


class DeviceAtmosphere:
    """All inputs of an Atmosphere uploaded once (inputs resident in HBM)."""

    def __init__(self, atm, precision=8, cloudy=None):
        dt = real_dtype(precision)
        self.precision = precision
        self.ncol, self.nlay = atm.p_lay.shape
        self.top_at_1 = atm.top_at_1
        self.p_lay, self.p_lev, self.t_lay, self.t_lev = (DeviceArray.from_host(a, dt) for a in (
            atm.p_lay, atm.p_lev, atm.t_lay, atm.t_lev))
        self.t_sfc = DeviceArray.from_host(atm.t_sfc, dt)
        self.emis_sfc = DeviceArray.from_host(atm.emis_sfc, dt)
        self.sfc_alb = DeviceArray.from_host(atm.sfc_alb, dt)
        self.mu0 = DeviceArray.from_host(atm.mu0, dt)
        self.gas_concs = GasConcs(atm.vmr, precision)
        self.cloudy = (atm.lwp is not None) if cloudy is None else cloudy
        if self.cloudy:
            self.lwp, self.iwp, self.rel, self.rei = (DeviceArray.from_host(a, dt) for a in (
                atm.lwp, atm.iwp, atm.rel, atm.rei))
        else:
            self.lwp = self.iwp = self.rel = self.rei = None


class SolverLW(_Solver):
    """Radiation_solver_longwave_gpu."""

    def __init__(self, kd: KDist, cloud: Cloud | None = None):
        self.precision = kd.precision
        self.kd, self.cloud = kd, cloud
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_solver_lw_create(kd.h, cloud.h if cloud else None, C.byref(h)))
        self.h = h
        self.out = None

    def solve(self, da: DeviceAtmosphere, bnd=False, col_dry=None, outputs=None, heating_rate=False):
        """Broadband (and with bnd, band) fluxes; with heating_rate, also o["heating_rate"]
        (ncol, nlay) [K/day] from the solver's FP64 sums of the layers' net-flux convergence."""
        dt = real_dtype(self.precision)
        ncol, nlay = da.ncol, da.nlay
        nlev = nlay + 1
        o = outputs
        if o is None:
            o = {k: DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_net")}
            for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_net"):
                o[k] = DeviceArray((ncol, nlev, self.kd.nbnd), dt) if bnd else None
        if heating_rate and o.get("heating_rate") is None:
            o["heating_rate"] = DeviceArray((ncol, nlay), dt)
        cloudy = da.cloudy and self.cloud is not None
        col_dry = _dev(col_dry, dt)
        self._check(self.L.rrtmgp_api_solve_lw(
            self.h, ncol, nlay, int(cloudy), int(bnd), da.p_lay.ptr, da.p_lev.ptr, da.t_lay.ptr,
            da.t_lev.ptr, da.t_sfc.ptr, da.gas_concs.h, _p(col_dry), da.emis_sfc.ptr,
            _p(da.lwp) if cloudy else None, _p(da.iwp) if cloudy else None,
            _p(da.rel) if cloudy else None, _p(da.rei) if cloudy else None,
            o["flux_up"].ptr, o["flux_dn"].ptr, o["flux_net"].ptr, _p(o.get("bnd_flux_up")),
            _p(o.get("bnd_flux_dn")), _p(o.get("bnd_flux_net")), _p(o.get("heating_rate"))))
        return o


    def adjoint(self, da: DeviceAtmosphere, g_flux_up, g_flux_dn):
        """Gradients of sum(g_flux_up * flux_up + g_flux_dn * flux_dn) over the broadband fluxes
        (ncol, nlev) of solve(da) (rrtmgp_api_adjoint_lw).  Returns device arrays: t_lay, t_lev,
        t_sfc, lnp_lay (with respect to ln p_lay), p_lev, vmr (ncol, nlay, ngas; the k-distribution's
        gas order) and emis_sfc (nbnd, ncol)."""
        dt = real_dtype(self.precision)
        ncol, nlay = da.ncol, da.nlay
        nlev = nlay + 1
        gu, gd = _dev(g_flux_up, dt), _dev(g_flux_dn, dt)
        o = {"t_lay": DeviceArray((ncol, nlay), dt), "t_lev": DeviceArray((ncol, nlev), dt),
             "t_sfc": DeviceArray((ncol,), dt), "lnp_lay": DeviceArray((ncol, nlay), dt),
             "p_lev": DeviceArray((ncol, nlev), dt), "vmr": DeviceArray((ncol, nlay, self.kd.ngas), dt),
             "emis_sfc": DeviceArray((self.kd.nbnd, ncol), dt)}
        cloudy = da.cloudy and self.cloud is not None
        self._check(self.L.rrtmgp_api_adjoint_lw(
            self.h, ncol, nlay, int(cloudy), da.p_lay.ptr, da.p_lev.ptr, da.t_lay.ptr, da.t_lev.ptr, da.t_sfc.ptr,
            da.gas_concs.h, da.emis_sfc.ptr, _p(da.lwp) if cloudy else None, _p(da.iwp) if cloudy else None,
            _p(da.rel) if cloudy else None, _p(da.rei) if cloudy else None, gu.ptr, gd.ptr,
            o["t_lay"].ptr, o["t_lev"].ptr, o["t_sfc"].ptr, o["lnp_lay"].ptr, o["p_lev"].ptr, o["vmr"].ptr,
            o["emis_sfc"].ptr))
        return o


class SolverSW(_Solver):
    """Radiation_solver_shortwave_gpu."""

    def __init__(self, kd: KDist, cloud: Cloud | None = None, aerosol: Aerosol | None = None):
        self.precision = kd.precision
        self.kd, self.cloud, self.aerosol = kd, cloud, aerosol
        h = C.c_void_p()
        self._check(self.L.rrtmgp_api_solver_sw_create_aerosol(kd.h, cloud.h if cloud else None,
                                                               aerosol.h if aerosol else None, C.byref(h)))
        self.h = h

    def solve(self, da: DeviceAtmosphere, tsi_scaling, bnd=False, delta_cloud=True, col_dry=None, outputs=None,
              heating_rate=False, aerosol=None, delta_aerosol=False, sfc_alb_dif=None):
        """As SolverLW.solve, with the direct flux; heating_rate as there.  aerosol = (rh, concs):
        relative humidity (ncol, nlay) and a GasConcs of aermr01..aermr11, added after the clouds
        (delta-scaled with delta_aerosol).  The diffuse albedo (nbnd, ncol) defaults to the direct
        one (da.sfc_alb)."""
        dt = real_dtype(self.precision)
        ncol, nlay = da.ncol, da.nlay
        nlev = nlay + 1
        tsi = _dev(tsi_scaling, dt)
        o = outputs
        if o is None:
            o = {k: DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net")}
            for k in ("bnd_flux_up", "bnd_flux_dn", "bnd_flux_dn_dir", "bnd_flux_net"):
                o[k] = DeviceArray((ncol, nlev, self.kd.nbnd), dt) if bnd else None
        if heating_rate and o.get("heating_rate") is None:
            o["heating_rate"] = DeviceArray((ncol, nlay), dt)
        cloudy = da.cloudy and self.cloud is not None
        rh, aconcs = (None, None) if aerosol is None else aerosol
        rh = _dev(rh, dt) if aerosol is not None else None
        col_dry = _dev(col_dry, dt)
        af = da.sfc_alb if sfc_alb_dif is None else _dev(sfc_alb_dif, dt)
        self._check(self.L.rrtmgp_api_solve_sw_aerosol(
            self.h, ncol, nlay, int(cloudy), int(aerosol is not None), int(bnd), int(delta_cloud),
            int(delta_aerosol), da.p_lay.ptr, da.p_lev.ptr,
            da.t_lay.ptr, da.t_lev.ptr, da.gas_concs.h, _p(col_dry), da.sfc_alb.ptr, af.ptr,
            tsi.ptr, da.mu0.ptr, _p(da.lwp) if cloudy else None, _p(da.iwp) if cloudy else None,
            _p(da.rel) if cloudy else None, _p(da.rei) if cloudy else None,
            _p(rh), aconcs.h if aconcs is not None else None,
            o["flux_up"].ptr, o["flux_dn"].ptr, o["flux_dn_dir"].ptr, o["flux_net"].ptr,
            _p(o.get("bnd_flux_up")), _p(o.get("bnd_flux_dn")), _p(o.get("bnd_flux_dn_dir")),
            _p(o.get("bnd_flux_net")), _p(o.get("heating_rate"))))
        return o

    def adjoint(self, da: DeviceAtmosphere, tsi_scaling, g_flux_up, g_flux_dn, g_flux_dir, delta_cloud=True,
                sfc_alb_dif=None):
        """Gradients of sum(g_flux_up * flux_up + g_flux_dn * flux_dn + g_flux_dir * flux_dn_dir) over
        the broadband fluxes of solve(da, tsi_scaling) (rrtmgp_api_adjoint_sw).  The diffuse albedo
        defaults to the direct one, as in solve().  Returns device arrays: t_lay, lnp_lay, p_lev,
        vmr (ncol, nlay, ngas), sfc_alb_dir and sfc_alb_dif (nbnd, ncol)."""
        dt = real_dtype(self.precision)
        ncol, nlay = da.ncol, da.nlay
        nlev = nlay + 1
        tsi = _dev(tsi_scaling, dt)
        af = da.sfc_alb if sfc_alb_dif is None else _dev(sfc_alb_dif, dt)
        gu, gd, gr = _dev(g_flux_up, dt), _dev(g_flux_dn, dt), _dev(g_flux_dir, dt)
        o = {"t_lay": DeviceArray((ncol, nlay), dt), "lnp_lay": DeviceArray((ncol, nlay), dt),
             "p_lev": DeviceArray((ncol, nlev), dt), "vmr": DeviceArray((ncol, nlay, self.kd.ngas), dt),
             "sfc_alb_dir": DeviceArray((self.kd.nbnd, ncol), dt),
             "sfc_alb_dif": DeviceArray((self.kd.nbnd, ncol), dt)}
        cloudy = da.cloudy and self.cloud is not None
        self._check(self.L.rrtmgp_api_adjoint_sw(
            self.h, ncol, nlay, int(cloudy), int(delta_cloud), da.p_lay.ptr, da.p_lev.ptr, da.t_lay.ptr,
            da.gas_concs.h, da.sfc_alb.ptr, af.ptr, tsi.ptr, da.mu0.ptr,
            _p(da.lwp) if cloudy else None, _p(da.iwp) if cloudy else None,
            _p(da.rel) if cloudy else None, _p(da.rei) if cloudy else None, gu.ptr, gd.ptr, gr.ptr,
            o["t_lay"].ptr, o["lnp_lay"].ptr, o["p_lev"].ptr, o["vmr"].ptr, o["sfc_alb_dir"].ptr,
            o["sfc_alb_dif"].ptr))
        return o
