"""Seeded synthetic RRTMGP inputs: k-distribution tables, cloud lookup tables and profiles.

The real coefficient files (``rrtmgp-gas-lw-g256.nc`` and friends) are not available offline, so
the workloads are drawn from a seed.  Every array is returned in the layout the reference's
coefficient reader hands to the ``Gas_optics_rrtmgp`` constructors: Fortran order, first index
fastest, with the dimension order of ``src_test/Radiation_solver.cpp:70-284``
(``load_and_init_gas_optics``).  E.g. ``kmajor`` has Fortran dims (ngpt, neta, npres+1, ntemp)
which the constructor reorders to (ntemp, neta, npres+1, ngpt)
(``src/Gas_optics_rrtmgp.cpp:658-663``).

Generator choices that the workload description leaves open are fixed here and listed in
``GENERATOR_NOTES``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import log_ndtr

GENERATOR_NOTES = """
* k-table magnitude: ln k ~ N(-8, 3^2) at every table node as specified, multiplied by a
  reference cross-section K_SCALE (2e-20 cm^2 longwave, 1e-21 cm^2 shortwave).  The nodes of
  one g-point are not drawn independently: ln k = -8 + 3 (sqrt(1/2) a_g + sqrt(1/2) s_g) with
  a_g ~ N(0, 1) per g-point and s_g(eta, p, T) a smooth Gaussian field with exactly unit
  variance at every node (multilinear interpolation of an iid N(0, 1) coarse grid of 3 x 8 x 4
  nodes over (eta, p, T), each interpolation row scaled to unit L2 norm).  Real tables vary
  smoothly between neighbouring reference temperatures, pressures and eta values; independent
  draws would make neighbours differ by up to ~1e9, far from any physical table, and would turn
  the rounding of the interpolation weights into O(1) relative errors at any finite precision.
  Minor tables use the same construction over (eta, T) with a 3 x 4 coarse grid.  Taken
  literally (k ~ e^-8 cm^2 per molecule) every layer would have an optical depth of order 1e18;
  with the scales the median clear-sky column optical depth of the nlay=60 profiles is ~2.9
  (LW) / ~0.4 (SW), the mean OLR is ~230 W m-2 and ~2/3 of the TOA solar flux reaches the
  surface.
* gases known to the k-distribution: h2o, co2, o3, n2o, ch4.  ch4 only ever appears as a minor
  contributor, so dropping it from the available gases exercises the minor-table reduction.
* Planck fractions: Dirichlet(1) over the g-points of a band at every (eta, p, T) node, made
  smooth across nodes by a Gaussian copula: the Gamma(1) variates are -ln(1 - Phi(s)) of smooth
  unit-variance fields s built as for the k-tables.
* key species per band are fixed lists (LW_KEY_LOWER/UPPER, SW_KEY_LOWER/UPPER).
* press_ref_trop = 9948.431564193395 Pa; vmr_ref = typical mixing ratios x exp(U(-0.3, 0.3))
  per (tropo half, reference temperature); "dry air" (gas 0) reference vmr = 1.
* minor intervals each cover the full g-point range of one band; 21 lower and 16 upper
  intervals; kminor drawn like kmajor.
* totplnk: band-integrated Planck radiance (W m-2 sr-1) on 196 temperatures 160..355 K
  (1 K step), integrated numerically over the band wavenumber limits.
* solar source: 5778 K Planck spectrum scaled to 1360 W m-2 in total, split over the g-points of
  a band with Dirichlet(1) weights; facular/sunspot terms are +-1e-3 of the quiet term and the
  default mg/sb indices equal the offsets + 0.01.
* Rayleigh: 4.1e-28 cm^2 * (nu / 1e4 cm-1)^4 at band centre, jitter exp(N(0, 0.1)).
* cloud LUTs: 20 sizes per band; liquid r_eff 2.5..21.5 um, ice D_eff 10..180 um; extinction
  1.5/r_eff (m2/g, liquid) or 3/D_eff (ice) times U(0.8, 1.2); ssa U(0.5, 1), g U(0.7, 0.9).
* profiles: level z = 7 km * ln(1.013e5 / p); play = geometric mean of the bounding levels;
  o3 Gaussian in ln p with sigma 1.0 around 1000 Pa on a 3e-8 background (without it o3 falls
  to ~1e-16 at 1 Pa, far below any observed value); level order is surface first
  (top_at_1 = False) unless top_at_1=True is asked for.
* ice clouds: same 20% layer fraction as liquid.
"""

O3_BACKGROUND = 3.0e-8
K_SCALE = {True: 2.0e-20, False: 1.0e-21}   # longwave, shortwave
GAS_NAMES = ["h2o", "co2", "o3", "n2o", "ch4"]

# Real RRTMGP band limits (cm-1), a physical convention, not data of the reference.
LW_BAND_EDGES = [10., 250., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480.,
                 1800., 2080., 2250., 2390., 2680., 3250.]
SW_BAND_EDGES = [820., 2680., 3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850.,
                 16000., 22650., 29000., 38000., 50000.]

# (gas1, gas2) 1-based indices into GAS_NAMES, 0 = none; (0, 0) is rewritten to (2, 2)
# by create_flavor (src/Gas_optics_rrtmgp.cpp:261-268).
LW_KEY_LOWER = [(1, 0), (1, 2), (1, 2), (1, 2), (1, 2), (1, 3), (1, 2), (1, 3),
                (1, 4), (1, 4), (1, 4), (1, 2), (1, 2), (4, 2), (1, 2), (1, 2)]
LW_KEY_UPPER = [(1, 0), (1, 2), (2, 3), (2, 3), (2, 3), (3, 2), (3, 0), (3, 2),
                (1, 3), (3, 0), (1, 4), (1, 2), (2, 3), (4, 2), (1, 2), (0, 0)]
SW_KEY_LOWER = [(1, 2), (1, 2), (1, 2), (1, 2), (1, 0), (1, 2), (1, 0), (1, 2),
                (1, 3), (1, 0), (1, 3), (3, 0), (3, 0), (3, 0)]
SW_KEY_UPPER = [(1, 2), (2, 0), (2, 0), (1, 2), (1, 0), (1, 2), (1, 0), (2, 0),
                (3, 2), (1, 0), (3, 0), (3, 0), (3, 0), (0, 0)]

IDENTIFIER_MINOR = ["h2o_frgn", "h2o_self", "co2", "o3", "n2o", "ch4"]
GAS_MINOR = ["h2o", "h2o", "co2", "o3", "n2o", "ch4"]

PRESS_REF_TROP = 9948.431564193395
NTEMP, NETA, NPRES = 14, 9, 59
N_PLANCK_TEMP = 196

H_PLANCK = 6.62607015e-34
C_LIGHT = 2.99792458e8
K_BOLTZ = 1.380649e-23


def fa(x, dtype=np.float64):
    """Fortran-ordered copy with the requested dtype."""
    return np.asfortranarray(np.asarray(x, dtype=dtype))


def planck_band_radiance(t, nu1, nu2, n=400):
    """Band-integrated Planck radiance (W m-2 sr-1) between nu1 and nu2 (cm-1), Simpson rule."""
    nu = np.linspace(nu1, nu2, 2 * n + 1) * 100.0  # m-1
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    x = H_PLANCK * C_LIGHT * nu[None, :] / (K_BOLTZ * t[:, None])
    b = 2.0 * H_PLANCK * C_LIGHT ** 2 * nu[None, :] ** 3 / np.expm1(x)
    h = (nu[1] - nu[0])
    w = np.ones(nu.size)
    w[1:-1:2] = 4.0
    w[2:-1:2] = 2.0
    return (b * w[None, :]).sum(axis=1) * h / 3.0


@dataclass
class KDist:
    """A k-distribution in the reference's coefficient-file layout (Fortran dims)."""
    is_lw: bool
    gas_names: list
    key_species: np.ndarray          # int (2, 2, nbnd)
    band2gpt: np.ndarray             # int (2, nbnd), 1-based inclusive
    band_lims_wavenum: np.ndarray    # (2, nbnd)
    press_ref: np.ndarray            # (npres)
    press_ref_trop: float
    temp_ref: np.ndarray             # (ntemp)
    temp_ref_p: float
    temp_ref_t: float
    vmr_ref: np.ndarray              # (2, ngas+1, ntemp)
    kmajor: np.ndarray               # (ngpt, neta, npres+1, ntemp)
    gas_minor: list
    identifier_minor: list
    kminor_lower: np.ndarray         # (ncontrib, neta, ntemp)
    kminor_upper: np.ndarray
    minor_gases_lower: list
    minor_gases_upper: list
    minor_limits_gpt_lower: np.ndarray   # int (2, nminor)
    minor_limits_gpt_upper: np.ndarray
    minor_scales_with_density_lower: np.ndarray  # int8 (nminor)
    minor_scales_with_density_upper: np.ndarray
    scaling_gas_lower: list
    scaling_gas_upper: list
    scale_by_complement_lower: np.ndarray        # int8 (nminor)
    scale_by_complement_upper: np.ndarray
    kminor_start_lower: np.ndarray               # int (nminor), 1-based
    kminor_start_upper: np.ndarray
    totplnk: np.ndarray | None = None            # (nPlanckTemp, nbnd)
    planck_frac: np.ndarray | None = None        # (ngpt, neta, npres+1, ntemp)
    solar_src_quiet: np.ndarray | None = None    # (ngpt)
    solar_src_facular: np.ndarray | None = None
    solar_src_sunspot: np.ndarray | None = None
    tsi_default: float = 0.0
    mg_default: float = 0.0
    sb_default: float = 0.0
    rayl_lower: np.ndarray | None = None         # (ngpt, neta, ntemp)
    rayl_upper: np.ndarray | None = None

    @property
    def nbnd(self):
        return self.band2gpt.shape[1]

    @property
    def ngpt(self):
        return int(self.band2gpt.max())


def _interp_rows(n_fine, n_coarse):
    """(n_fine, n_coarse) linear-interpolation matrix, each row scaled to unit L2 norm."""
    x = np.linspace(0.0, n_coarse - 1.0, n_fine)
    i0 = np.minimum(np.floor(x).astype(int), n_coarse - 2)
    w = x - i0
    a = np.zeros((n_fine, n_coarse))
    a[np.arange(n_fine), i0] = 1.0 - w
    a[np.arange(n_fine), i0 + 1] = w
    return a / np.sqrt((a * a).sum(axis=1, keepdims=True))


def _smooth_unit_field(rng, nlead, fine, coarse):
    """nlead independent Gaussian fields on the node grid ``fine`` (one axis per entry), each
    N(0, 1) at every node and smooth between nodes: multilinear interpolation of an iid N(0, 1)
    grid of ``coarse`` nodes with row-normalised weights (the tensor product of unit-norm rows
    has unit norm, so every node keeps unit variance).  Returns shape (nlead, *fine)."""
    f = rng.normal(size=(nlead, *coarse))
    for ax, (nf, nc) in enumerate(zip(fine, coarse)):
        f = np.moveaxis(np.tensordot(f, _interp_rows(nf, nc), axes=([ax + 1], [1])), -1, ax + 1)
    return f


def _lognormal_table(rng, nlead, fine, coarse, scale):
    """exp(-8 + 3 z) * scale with z = sqrt(1/2) a + sqrt(1/2) s: N(0, 1) at every node."""
    a = rng.normal(size=(nlead,) + (1,) * len(fine))
    z = math.sqrt(0.5) * a + math.sqrt(0.5) * _smooth_unit_field(rng, nlead, fine, coarse)
    return np.exp(-8.0 + 3.0 * z) * scale


def _minor_block(rng, nbnd, gpt_per_band, n_intervals, first_bands, scale):
    """Minor intervals: one per band in ``first_bands`` order, extra ones on random bands."""
    bands = list(first_bands)
    extra = n_intervals - len(bands)
    if extra > 0:
        bands += list(rng.choice(nbnd, size=extra, replace=False))
    bands = sorted(bands)
    idents, dens, scal, compl = [], [], [], []
    used = {}
    for b in bands:
        choices = [i for i in IDENTIFIER_MINOR if i not in used.get(b, set())]
        ident = choices[rng.integers(len(choices))]
        used.setdefault(b, set()).add(ident)
        idents.append(ident)
        if ident == "h2o_self":
            dens.append(1); scal.append("h2o"); compl.append(0)
        elif ident == "h2o_frgn":
            dens.append(1); scal.append("h2o"); compl.append(1)
        else:
            d = int(rng.integers(2))
            dens.append(d)
            # A few density-scaled contributors use a scaling gas, as in the real tables.
            if d == 1 and rng.random() < 0.3:
                scal.append("co2" if ident != "co2" else "h2o")
                compl.append(int(rng.integers(2)))
            else:
                scal.append("")
                compl.append(0)
    limits = np.zeros((2, len(bands)), dtype=np.int32)
    start = np.zeros(len(bands), dtype=np.int32)
    pos = 1
    for i, b in enumerate(bands):
        limits[0, i] = b * gpt_per_band + 1
        limits[1, i] = (b + 1) * gpt_per_band
        start[i] = pos
        pos += gpt_per_band
    ncontrib = pos - 1
    kminor = _lognormal_table(rng, ncontrib, (NETA, NTEMP), (3, 4), scale)
    return (fa(kminor), idents, fa(limits, np.int32), np.array(dens, dtype=np.int8), scal,
            np.array(compl, dtype=np.int8), np.array(start, dtype=np.int32))


def make_kdist(is_lw: bool, seed: int) -> KDist:
    rng = np.random.Generator(np.random.PCG64(seed))
    edges = LW_BAND_EDGES if is_lw else SW_BAND_EDGES
    nbnd = len(edges) - 1
    gpb = 16
    ngpt = nbnd * gpb
    band2gpt = np.zeros((2, nbnd), dtype=np.int32)
    band2gpt[0] = np.arange(nbnd) * gpb + 1
    band2gpt[1] = (np.arange(nbnd) + 1) * gpb
    band_lims = np.array([edges[:-1], edges[1:]], dtype=np.float64)

    klow = LW_KEY_LOWER if is_lw else SW_KEY_LOWER
    kupp = LW_KEY_UPPER if is_lw else SW_KEY_UPPER
    key_species = np.zeros((2, 2, nbnd), dtype=np.int32)
    for b in range(nbnd):
        key_species[:, 0, b] = klow[b]
        key_species[:, 1, b] = kupp[b]

    press_ref = np.exp(np.linspace(math.log(1.1e5), math.log(1.0), NPRES))
    temp_ref = np.linspace(160.0, 355.0, NTEMP)

    typical = {"h2o": (5e-3, 4e-6), "co2": (3.5e-4, 3.5e-4), "o3": (5e-8, 3e-6),
               "n2o": (3e-7, 2e-7), "ch4": (1.7e-6, 1.2e-6)}
    ngas = len(GAS_NAMES)
    vmr_ref = np.zeros((2, ngas + 1, NTEMP))
    vmr_ref[:, 0, :] = 1.0
    for ig, g in enumerate(GAS_NAMES):
        for itr in range(2):
            vmr_ref[itr, ig + 1, :] = typical[g][itr] * np.exp(rng.uniform(-0.3, 0.3, NTEMP))

    kmajor = _lognormal_table(rng, ngpt, (NETA, NPRES + 1, NTEMP), (3, 8, 4), K_SCALE[is_lw])

    n_low, n_upp = 21, 16
    lo = _minor_block(rng, nbnd, gpb, n_low, range(min(nbnd, n_low)), K_SCALE[is_lw])
    up = _minor_block(rng, nbnd, gpb, n_upp, range(min(nbnd, n_upp)), K_SCALE[is_lw])

    kd = KDist(
        is_lw=is_lw, gas_names=list(GAS_NAMES), key_species=fa(key_species, np.int32),
        band2gpt=fa(band2gpt, np.int32), band_lims_wavenum=fa(band_lims),
        press_ref=fa(press_ref), press_ref_trop=PRESS_REF_TROP, temp_ref=fa(temp_ref),
        temp_ref_p=100.0 * 100.0, temp_ref_t=250.0, vmr_ref=fa(vmr_ref), kmajor=fa(kmajor),
        gas_minor=list(GAS_MINOR), identifier_minor=list(IDENTIFIER_MINOR),
        kminor_lower=lo[0], minor_gases_lower=lo[1], minor_limits_gpt_lower=lo[2],
        minor_scales_with_density_lower=lo[3], scaling_gas_lower=lo[4],
        scale_by_complement_lower=lo[5], kminor_start_lower=lo[6],
        kminor_upper=up[0], minor_gases_upper=up[1], minor_limits_gpt_upper=up[2],
        minor_scales_with_density_upper=up[3], scaling_gas_upper=up[4],
        scale_by_complement_upper=up[5], kminor_start_upper=up[6])

    if is_lw:
        tplanck = np.linspace(160.0, 355.0, N_PLANCK_TEMP)
        totplnk = np.zeros((N_PLANCK_TEMP, nbnd))
        for b in range(nbnd):
            totplnk[:, b] = planck_band_radiance(tplanck, edges[b], edges[b + 1])
        kd.totplnk = fa(totplnk)
        # Gamma(1) = Exp(1) variates -ln(1 - Phi(s)) = -ln Phi(-s) of smooth N(0, 1) fields.
        pf = -log_ndtr(-_smooth_unit_field(rng, ngpt, (NETA, NPRES + 1, NTEMP), (3, 8, 4)))
        pf = pf.reshape(nbnd, gpb, NETA, NPRES + 1, NTEMP)
        pf /= pf.sum(axis=1, keepdims=True)
        kd.planck_frac = fa(pf.reshape(ngpt, NETA, NPRES + 1, NTEMP))
    else:
        band_sun = np.array([planck_band_radiance(5778.0, edges[b], edges[b + 1])[0]
                             for b in range(nbnd)])
        band_sun *= 1360.0 / band_sun.sum()
        w = rng.gamma(1.0, size=(nbnd, gpb))
        w /= w.sum(axis=1, keepdims=True)
        quiet = (band_sun[:, None] * w).reshape(ngpt)
        kd.solar_src_quiet = fa(quiet)
        kd.solar_src_facular = fa(quiet * 1e-3 * rng.uniform(-1, 1, ngpt))
        kd.solar_src_sunspot = fa(quiet * 1e-3 * rng.uniform(-1, 1, ngpt))
        kd.tsi_default = 1360.0
        kd.mg_default = 0.1495954 + 0.01
        kd.sb_default = 0.00066696 + 0.01
        centre = 0.5 * (band_lims[0] + band_lims[1])
        ray = np.zeros((2, ngpt, NETA, NTEMP))
        for b in range(nbnd):
            base = 4.1e-28 * (centre[b] / 1.0e4) ** 4
            ray[:, b * gpb:(b + 1) * gpb] = base * np.exp(
                rng.normal(0.0, 0.1, size=(2, gpb, NETA, NTEMP)))
        kd.rayl_lower = fa(ray[0])
        kd.rayl_upper = fa(ray[1])
    return kd


@dataclass
class CloudLUT:
    """Cloud optics lookup tables in the reference's constructor layout
    (``src/Cloud_optics.cpp:29-69``): liquid (nsize, nbnd), ice (nsize, nbnd, nrough=3)."""
    band_lims_wvn: np.ndarray
    radliq_lwr: float
    radliq_upr: float
    radliq_fac: float
    radice_lwr: float
    radice_upr: float
    radice_fac: float
    lut_extliq: np.ndarray
    lut_ssaliq: np.ndarray
    lut_asyliq: np.ndarray
    lut_extice: np.ndarray
    lut_ssaice: np.ndarray
    lut_asyice: np.ndarray

    @property
    def nbnd(self):
        return self.band_lims_wvn.shape[1]


def make_cloud_lut(is_lw: bool, seed: int, nsize: int = 20) -> CloudLUT:
    rng = np.random.Generator(np.random.PCG64(seed))
    edges = LW_BAND_EDGES if is_lw else SW_BAND_EDGES
    nbnd = len(edges) - 1
    rl = np.linspace(2.5, 21.5, nsize)
    ri = np.linspace(10.0, 180.0, nsize)
    extliq = (1.5 / rl)[:, None] * rng.uniform(0.8, 1.2, (nsize, nbnd))
    extice = (3.0 / ri)[:, None, None] * rng.uniform(0.8, 1.2, (nsize, nbnd, 3))
    return CloudLUT(
        band_lims_wvn=fa(np.array([edges[:-1], edges[1:]])),
        radliq_lwr=2.5, radliq_upr=21.5, radliq_fac=1.0,
        radice_lwr=10.0, radice_upr=180.0, radice_fac=1.0,
        lut_extliq=fa(extliq), lut_ssaliq=fa(rng.uniform(0.5, 1.0, (nsize, nbnd))),
        lut_asyliq=fa(rng.uniform(0.7, 0.9, (nsize, nbnd))),
        lut_extice=fa(extice), lut_ssaice=fa(rng.uniform(0.5, 1.0, (nsize, nbnd, 3))),
        lut_asyice=fa(rng.uniform(0.7, 0.9, (nsize, nbnd, 3))))


@dataclass
class Atmosphere:
    """Column profiles in the reference's driver layout (``test_rte_rrtmgp.cpp``):
    (ncol, nlay|nlev) Fortran order; gas vmr are (ncol, nlay) or (1, 1) scalars."""
    p_lay: np.ndarray
    p_lev: np.ndarray
    t_lay: np.ndarray
    t_lev: np.ndarray
    t_sfc: np.ndarray
    vmr: dict
    emis_sfc: np.ndarray      # (nbnd_lw, ncol)
    sfc_alb: np.ndarray       # (nbnd_sw, ncol), used for both direct and diffuse
    mu0: np.ndarray           # (ncol)
    tsi: float
    lwp: np.ndarray | None = None
    iwp: np.ndarray | None = None
    rel: np.ndarray | None = None
    rei: np.ndarray | None = None
    top_at_1: bool = False
    extra: dict = field(default_factory=dict)

    @property
    def ncol(self):
        return self.p_lay.shape[0]

    @property
    def nlay(self):
        return self.p_lay.shape[1]


def make_atmosphere(seed: int, ncol: int, nlay: int, cloudy: bool = False,
                    top_at_1: bool = False, nbnd_lw: int = 16, nbnd_sw: int = 14,
                    conservative_frac: float = 0.0) -> Atmosphere:
    # Fields are drawn and computed as (nz, ncol) C-order arrays, whose transposes are the
    # (ncol, nz) Fortran-order arrays of the driver layout without a copy.
    rng = np.random.Generator(np.random.PCG64(seed))
    nlev = nlay + 1
    p0 = np.exp(np.linspace(math.log(1.013e5), math.log(1.0), nlev))
    p_lev = rng.uniform(0.99, 1.01, size=(nlev, ncol))
    p_lev *= p0[:, None]
    np.clip(p_lev, 1.0, 1.1e5, out=p_lev)
    p_lay = np.sqrt(p_lev[:-1] * p_lev[1:])

    h = 7.0
    t_sfc = rng.uniform(280.0, 310.0, size=ncol)
    z_trop = rng.uniform(10.0, 16.0, size=ncol)

    def temp(p):
        z = np.log(1.013e5 / p)
        z *= h
        np.clip(z, 0.0, z_trop[None, :], out=z)      # max(z, 0), isothermal above z_trop
        return t_sfc[None, :] - 6.5 * z

    t_lev = temp(p_lev)
    t_lay = temp(p_lay)
    t_sfc_out = t_lev[0].copy()

    z_lay = np.maximum(h * np.log(1.013e5 / p_lay), 0.0)
    h2o = np.maximum(1e-2 * np.exp(-z_lay / 2.0), 3e-6)
    lnp = np.log(p_lay) - math.log(1000.0)
    o3 = 8e-6 * np.exp(-0.5 * lnp * lnp) + O3_BACKGROUND
    del z_lay, lnp
    vmr = {"h2o": h2o.T, "o3": o3.T, "co2": fa([[400e-6]]), "ch4": fa([[1.8e-6]]),
           "n2o": fa([[3.2e-7]])}

    emis = np.full((nbnd_lw, ncol), 0.98, order="F")
    alb = fa(np.repeat(rng.uniform(0.06, 0.3, size=ncol)[None, :], nbnd_sw, axis=0))
    mu0 = rng.uniform(0.1, 1.0, size=ncol)

    atm = Atmosphere(p_lay=p_lay.T, p_lev=p_lev.T, t_lay=t_lay.T, t_lev=t_lev.T, t_sfc=t_sfc_out,
                     vmr=vmr, emis_sfc=emis, sfc_alb=alb, mu0=mu0, tsi=1360.0)
    if cloudy:
        liq_zone = (p_lay <= 900e2) & (p_lay >= 400e2)
        ice_zone = (p_lay <= 400e2) & (p_lay >= 150e2)
        liq = liq_zone & (rng.random((nlay, ncol)) < 0.2)
        ice = ice_zone & (rng.random((nlay, ncol)) < 0.2)
        lwp = np.where(liq, 50.0 * np.exp(rng.normal(0.0, 1.0, (nlay, ncol))), 0.0)
        iwp = np.where(ice, 20.0 * np.exp(rng.normal(0.0, 1.0, (nlay, ncol))), 0.0)
        rel = np.where(liq, rng.uniform(4.0, 20.0, (nlay, ncol)), 0.0)
        rei = np.where(ice, rng.uniform(20.0, 120.0, (nlay, ncol)), 0.0)
        atm.lwp, atm.iwp, atm.rel, atm.rei = lwp.T, iwp.T, rel.T, rei.T
        if conservative_frac > 0.0:
            cl = liq | ice
            atm.extra["conservative"] = fa((cl & (rng.random((nlay, ncol)) < conservative_frac)).T)
    if top_at_1:
        atm = flip_vertical(atm)
    return atm


def flip_vertical(atm: Atmosphere) -> Atmosphere:
    """Reverse the vertical index (surface-first <-> top-first)."""
    def flip(a):
        return None if a is None else fa(a[:, ::-1])
    vmr = {k: (v if v.shape == (1, 1) else flip(v)) for k, v in atm.vmr.items()}
    extra = {k: (flip(v) if isinstance(v, np.ndarray) and v.ndim == 2 else v)
             for k, v in atm.extra.items()}
    return Atmosphere(p_lay=flip(atm.p_lay), p_lev=flip(atm.p_lev), t_lay=flip(atm.t_lay),
                      t_lev=flip(atm.t_lev), t_sfc=atm.t_sfc, vmr=vmr, emis_sfc=atm.emis_sfc,
                      sfc_alb=atm.sfc_alb, mu0=atm.mu0, tsi=atm.tsi, lwp=flip(atm.lwp),
                      iwp=flip(atm.iwp), rel=flip(atm.rel), rei=flip(atm.rei),
                      top_at_1=not atm.top_at_1, extra=extra)


def subset_columns(atm: Atmosphere, cols) -> Atmosphere:
    """The columns ``cols`` (an index array or slice) of an atmosphere."""
    def sub(a, axis=0):
        if a is None:
            return None
        return fa(np.take(a, np.arange(atm.ncol)[cols], axis=axis))
    vmr = {k: (v if v.shape[0] == 1 else sub(v)) for k, v in atm.vmr.items()}
    extra = {k: (sub(v) if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == atm.ncol else v)
             for k, v in atm.extra.items()}
    return Atmosphere(p_lay=sub(atm.p_lay), p_lev=sub(atm.p_lev), t_lay=sub(atm.t_lay),
                      t_lev=sub(atm.t_lev), t_sfc=sub(atm.t_sfc), vmr=vmr,
                      emis_sfc=sub(atm.emis_sfc, 1), sfc_alb=sub(atm.sfc_alb, 1), mu0=sub(atm.mu0),
                      tsi=atm.tsi, lwp=sub(atm.lwp), iwp=sub(atm.iwp), rel=sub(atm.rel),
                      rei=sub(atm.rei), top_at_1=atm.top_at_1, extra=extra)

# Seeds of the named workloads.
SEED_KDIST_LW = 20260101
SEED_KDIST_SW = 20260102
SEED_CLOUD_LW = 20260103
SEED_CLOUD_SW = 20260104
SEED_ATMOS = {0: 1000, 1: 1001, 2: 1002, 3: 1003, 4: 1004}
