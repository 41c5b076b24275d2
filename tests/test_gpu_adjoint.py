"""GPU adjoints against central finite differences of the FP64 CPU oracle.

The solver adjoints return the gradient of L = <a_up, broadband_up> + <a_dn, broadband_dn> for
random cotangents a_up, a_dn; each check compares <gradient, v> with the central difference
(L(x + h v) - L(x - h v)) / 2h of the oracle's FP64 forward along random directions v.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

# Gaussian secants and weights of Rte_lw::rte_lw (src/Rte_lw.cpp:100-113).
GAUSS = {1: ([1.66], [0.5]), 2: ([1.18350343, 2.81649655], [0.3180413817, 0.1819586183])}


@pytest.fixture(scope="module")
def G():
    from rrtmgp_amd import api, device
    if device.device_count() < 1:
        pytest.fail("no HIP device visible")
    return api


def _host(d):
    return d.to_host().astype(np.float64)


def _lw_case(O, okd_lw, syn, ncol, nlay, top_at_1, seed):
    atm = syn.make_atmosphere(seed, ncol, nlay, top_at_1=top_at_1)
    go = O.gas_optics(okd_lw, atm)
    bl = np.asarray(okd_lw.band_lims)
    band_of = np.concatenate([np.full(bl[1, b] - bl[0, b] + 1, b) for b in range(bl.shape[1])])
    return atm, go, bl, band_of


def _lw_adjoint(G, go, emis_gpt, a_up, a_dn, top_at_1, nmu, prec):
    from rrtmgp_amd import _lib
    dt = np.float64 if prec == 8 else np.float32
    ncol, nlay, ngpt = go["tau"].shape
    ds, ws = GAUSS[nmu]
    sec = np.empty((ncol, ngpt, nmu))
    for mu in range(nmu):
        sec[:, :, mu] = ds[mu]
    dev = lambda a: G.DeviceArray.from_host(np.asarray(a, dt))
    names = ("tau", "lay_src", "lev_src_inc", "lev_src_dec")
    d = {n: dev(go[n]) for n in names}
    sfc, emis, au, ad, dsec, dw = (dev(a) for a in (go["sfc_src"], emis_gpt, a_up, a_dn, sec, ws))
    out = {n: G.DeviceArray((ncol, nlay, ngpt), dt) for n in names}
    out["sfc_src"], out["emis"] = G.DeviceArray((ncol, ngpt), dt), G.DeviceArray((ncol, ngpt), dt)
    fn = getattr(_lib.hip(), "rrtmgp_hip_lw_solver_noscat_adjoint_" + ("f64" if prec == 8 else "f32"))
    _lib.check_hip(fn(C.c_int64(ncol), C.c_int64(nlay), C.c_int64(ngpt), C.c_int(int(top_at_1)), C.c_int64(nmu),
                      dsec.ptr, dw.ptr, d["tau"].ptr, d["lay_src"].ptr, d["lev_src_inc"].ptr, d["lev_src_dec"].ptr,
                      emis.ptr, sfc.ptr, None, au.ptr, ad.ptr, out["tau"].ptr, out["lay_src"].ptr,
                      out["lev_src_inc"].ptr, out["lev_src_dec"].ptr, out["sfc_src"].ptr, out["emis"].ptr, None))
    return {k: _host(v) for k, v in out.items()}


@pytest.mark.parametrize("top_at_1", [False, True])
@pytest.mark.parametrize("nmu", [1, 2])
def test_lw_solver_adjoint_fd(G, O, okd_lw, syn, top_at_1, nmu):
    atm, go, bl, band_of = _lw_case(O, okd_lw, syn, 6, 24, top_at_1, 41)
    ncol, nlay, ngpt = go["tau"].shape
    rng = np.random.default_rng(3)
    a_up, a_dn = rng.normal(size=(ncol, nlay + 1)), rng.normal(size=(ncol, nlay + 1))
    emis_b = np.asarray(atm.emis_sfc, np.float64) * rng.uniform(0.9, 1.0, np.shape(atm.emis_sfc))
    emis_gpt = emis_b[band_of, :].T
    g = _lw_adjoint(G, go, emis_gpt, a_up, a_dn, top_at_1, nmu, 8)

    def loss(x):
        up, dn = O.rte_lw(x["tau"], x["lay_src"], x["lev_src_inc"], x["lev_src_dec"], x["sfc_src"], x["emis"],
                          bl, top_at_1, n_gauss_angles=nmu, do_broadband=True)
        return float(np.sum(a_up * up) + np.sum(a_dn * dn))

    base = {k: go[k] for k in ("tau", "lay_src", "lev_src_inc", "lev_src_dec", "sfc_src")}
    base["emis"] = emis_b
    for trial in range(3):
        v = {k: rng.normal(size=np.shape(a)) * np.abs(a) for k, a in base.items()}
        v["emis"] = rng.normal(size=emis_b.shape) * 0.01
        h = 1e-5
        plus = {k: base[k] + h * v[k] for k in base}
        minus = {k: base[k] - h * v[k] for k in base}
        fd = (loss(plus) - loss(minus)) / (2 * h)
        ad = sum(np.sum(g[k] * v[k]) for k in ("tau", "lay_src", "lev_src_inc", "lev_src_dec", "sfc_src"))
        ad += np.sum(g["emis"] * v["emis"][band_of, :].T)
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)


def test_lw_solver_adjoint_fp32_vs_fp64(G, O, okd_lw, syn):
    atm, go, bl, band_of = _lw_case(O, okd_lw, syn, 64, 60, True, 42)
    ncol, nlay, ngpt = go["tau"].shape
    rng = np.random.default_rng(5)
    a_up, a_dn = rng.normal(size=(ncol, nlay + 1)), rng.normal(size=(ncol, nlay + 1))
    emis_gpt = np.asarray(atm.emis_sfc)[band_of, :].T
    g8 = _lw_adjoint(G, go, emis_gpt, a_up, a_dn, True, 1, 8)
    g4 = _lw_adjoint(G, go, emis_gpt, a_up, a_dn, True, 1, 4)
    for k in g8:
        scale = np.abs(g8[k]).max()
        assert np.abs(g4[k] - g8[k]).max() <= 1e-4 * scale, k


def _sw_inputs(O, okd_sw, lut_sw, syn, ncol, nlay, top_at_1, seed):
    atm = syn.make_atmosphere(seed, ncol, nlay, cloudy=True, top_at_1=top_at_1)
    go = O.gas_optics(okd_sw, atm)
    ct, cs, cg = O.delta_scale(*O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    tau, ssa, g = O.inc_bybnd(go["tau"], ct, okd_sw.band_lims, go["ssa"], go["g"], cs, cg)
    bl = np.asarray(okd_sw.band_lims)
    band_of = np.concatenate([np.full(bl[1, b] - bl[0, b] + 1, b) for b in range(bl.shape[1])])
    return atm, go, tau, ssa, g, bl, band_of


def _sw_adjoint(G, tau, ssa, g, mu0, inc, ad_gpt, af_gpt, a_up, a_dn, a_dir, top_at_1, prec):
    from rrtmgp_amd import _lib
    dt = np.float64 if prec == 8 else np.float32
    ncol, nlay, ngpt = tau.shape
    dev = lambda a: G.DeviceArray.from_host(np.asarray(a, dt))
    ins = [dev(a) for a in (tau, ssa, g, mu0, ad_gpt, af_gpt, inc)]
    cots = [dev(a) for a in (a_up, a_dn, a_dir)]
    out = [G.DeviceArray((ncol, nlay, ngpt), dt) for _ in range(3)] + [G.DeviceArray((ncol, ngpt), dt) for _ in range(2)]
    fn = getattr(_lib.hip(), "rrtmgp_hip_sw_solver_2stream_adjoint_" + ("f64" if prec == 8 else "f32"))
    _lib.check_hip(fn(C.c_int64(ncol), C.c_int64(nlay), C.c_int64(ngpt), C.c_int(int(top_at_1)),
                      *[a.ptr for a in ins[:6]], ins[6].ptr, None, *[a.ptr for a in cots],
                      *[a.ptr for a in out], None))
    return dict(zip(("tau", "ssa", "g", "alb_dir", "alb_dif"), (_host(a) for a in out)))


@pytest.mark.parametrize("top_at_1", [False, True])
def test_sw_solver_adjoint_fd(G, O, okd_sw, lut_sw, syn, top_at_1):
    atm, go, tau, ssa, g, bl, band_of = _sw_inputs(O, okd_sw, lut_sw, syn, 6, 24, top_at_1, 43)
    ncol, nlay, ngpt = tau.shape
    rng = np.random.default_rng(9)
    a_up, a_dn, a_dir = (rng.normal(size=(ncol, nlay + 1)) for _ in range(3))
    alb_dir = np.asarray(atm.sfc_alb, np.float64)
    alb_dif = 0.5 * alb_dir
    grad = _sw_adjoint(G, tau, ssa, g, atm.mu0, go["toa_src"], alb_dir[band_of, :].T, alb_dif[band_of, :].T,
                       a_up, a_dn, a_dir, top_at_1, 8)

    def loss(x):
        up, dn, dr = O.rte_sw(x["tau"], x["ssa"], x["g"], atm.mu0, go["toa_src"], x["alb_dir"], x["alb_dif"],
                              bl, top_at_1, do_broadband=True)
        return float(np.sum(a_up * up) + np.sum(a_dn * dn) + np.sum(a_dir * dr))

    base = {"tau": tau, "ssa": ssa, "g": g, "alb_dir": alb_dir, "alb_dif": alb_dif}
    for trial in range(3):
        v = {"tau": rng.normal(size=tau.shape) * tau,
             "ssa": rng.normal(size=ssa.shape) * np.minimum(ssa, 1 - ssa),
             "g": rng.normal(size=g.shape) * np.minimum(np.abs(g), 1 - np.abs(g)),
             "alb_dir": rng.normal(size=alb_dir.shape) * 0.01,
             "alb_dif": rng.normal(size=alb_dif.shape) * 0.01}
        h = 1e-5
        fd = (loss({k: base[k] + h * v[k] for k in base}) - loss({k: base[k] - h * v[k] for k in base})) / (2 * h)
        ad = sum(np.sum(grad[k] * v[k]) for k in ("tau", "ssa", "g"))
        ad += np.sum(grad["alb_dir"] * v["alb_dir"][band_of, :].T) + np.sum(grad["alb_dif"] * v["alb_dif"][band_of, :].T)
        assert abs(ad - fd) <= 1e-6 * abs(fd), (trial, ad, fd)


@pytest.mark.parametrize("conservative", [0.0, 0.1])
def test_sw_solver_adjoint_fp32_vs_fp64(G, O, okd_sw, lut_sw, syn, conservative):
    """The FP32 build's SW adjoint (FP32 inputs and state store, FP64 recursions) against the FP64
    build's on the same FP32-representable inputs, with and without near-conservative cloud
    layers (10% of the cloudy layers at ssa = 1 - 1e-7, g = 0.85 before delta-scaling)."""
    atm = syn.make_atmosphere(44, 48, 60, cloudy=True, conservative_frac=conservative)
    go = O.gas_optics(okd_sw, atm)
    ct, cs, cg = O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True)
    if conservative:
        cons = atm.extra["conservative"]
        assert cons.sum() > 0
        cs = np.where(cons[:, :, None], 1.0 - 1e-7, cs)
        cg = np.where(cons[:, :, None], 0.85, cg)
    ct, cs, cg = O.delta_scale(ct, cs, cg)
    tau, ssa, g = O.inc_bybnd(go["tau"], ct, okd_sw.band_lims, go["ssa"], go["g"], cs, cg)
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)          # noqa: E731
    tau, ssa, g = r32(tau), r32(ssa), r32(g)
    bl = np.asarray(okd_sw.band_lims)
    band_of = np.concatenate([np.full(bl[1, b] - bl[0, b] + 1, b) for b in range(bl.shape[1])])
    ncol, nlay, ngpt = tau.shape
    rng = np.random.default_rng(6)
    a_up, a_dn, a_dir = (r32(rng.normal(size=(ncol, nlay + 1))) for _ in range(3))
    alb = r32(np.asarray(atm.sfc_alb, np.float64)[band_of, :].T)
    args = (tau, ssa, g, r32(atm.mu0), r32(go["toa_src"]), alb, r32(0.5 * alb), a_up, a_dn, a_dir, False)
    g8 = _sw_adjoint(G, *args, 8)
    g4 = _sw_adjoint(G, *args, 4)
    errs = {k: float(np.abs(g4[k] - g8[k]).max() / np.abs(g8[k]).max()) for k in g8}
    print("SW adjoint FP32 vs FP64, max |diff| / max |FP64|:", errs)
    for k, e in errs.items():
        assert e <= 5e-6, (k, e)


# ---------------------------------------------------------------------------------------------
# End-to-end adjoints (rrtmgp_api_adjoint_{lw,sw}): band-by-band gas-optics recomputation, solver
# adjoint, gas-optics adjoint and the col_dry chain, against central differences of the FP64
# oracle's full all-sky solve along random directions in every input group.

def _perturbed(atm, v, h):
    import copy
    a = copy.deepcopy(atm)
    a.t_lay = atm.t_lay + h * v["t_lay"]
    a.p_lay = atm.p_lay * np.exp(h * v["lnp_lay"])
    a.p_lev = atm.p_lev + h * v["p_lev"]
    a.vmr = dict(atm.vmr)
    for gas in ("h2o", "o3"):
        a.vmr[gas] = atm.vmr[gas] + h * v[gas]
    if "t_lev" in v:
        a.t_lev = atm.t_lev + h * v["t_lev"]
        a.t_sfc = atm.t_sfc + h * v["t_sfc"]
        a.emis_sfc = atm.emis_sfc + h * v["emis_sfc"]
    if "sfc_alb" in v:
        a.sfc_alb = atm.sfc_alb + h * v["sfc_alb"]
    return a


def _directions(rng, atm, lw):
    v = {"t_lay": rng.normal(size=atm.t_lay.shape),
         # Pressures at the table's lower edge (1 Pa) stay put: the oracle's range check rejects
         # any step below it.
         "lnp_lay": rng.normal(size=atm.p_lay.shape) * 1e-3 * (atm.p_lay > 2.),
         "p_lev": rng.normal(size=atm.p_lev.shape) * 1e-3 * atm.p_lev * (atm.p_lev > 2.),
         "h2o": rng.normal(size=atm.vmr["h2o"].shape) * 0.01 * atm.vmr["h2o"],
         "o3": rng.normal(size=atm.vmr["o3"].shape) * 0.01 * atm.vmr["o3"]}
    if lw:
        v.update({"t_lev": rng.normal(size=atm.t_lev.shape), "t_sfc": rng.normal(size=atm.t_sfc.shape),
                  "emis_sfc": rng.normal(size=atm.emis_sfc.shape) * 0.01})
    else:
        v["sfc_alb"] = rng.normal(size=atm.sfc_alb.shape) * 0.01
    return v


def _dot(G, grad, v, names, lw):
    ih, io = names.index("h2o"), names.index("o3")
    vm = grad["vmr"]
    ad = (np.sum(grad["t_lay"] * v["t_lay"]) + np.sum(grad["lnp_lay"] * v["lnp_lay"]) +
          np.sum(grad["p_lev"] * v["p_lev"]) + np.sum(vm[:, :, ih] * v["h2o"]) + np.sum(vm[:, :, io] * v["o3"]))
    if lw:
        ad += (np.sum(grad["t_lev"] * v["t_lev"]) + np.sum(grad["t_sfc"] * v["t_sfc"]) +
               np.sum(grad["emis_sfc"] * v["emis_sfc"]))
    else:
        ad += np.sum((grad["sfc_alb_dir"] + grad["sfc_alb_dif"]) * v["sfc_alb"])
    return ad


@pytest.mark.parametrize("top_at_1", [False, True])
def test_lw_adjoint_end_to_end_fd(G, O, syn, kd_lw, okd_lw, lut_lw, top_at_1):
    from conftest import AVAILABLE
    atm = syn.make_atmosphere(51, 8, 30, cloudy=True, top_at_1=top_at_1)
    kd = G.KDist(kd_lw, AVAILABLE, precision=8)
    solver = G.SolverLW(kd, G.Cloud(lut_lw, 8))
    da = G.DeviceAtmosphere(atm, 8)
    rng = np.random.default_rng(11)
    a_up, a_dn = rng.normal(size=(atm.ncol, atm.nlay + 1)), rng.normal(size=(atm.ncol, atm.nlay + 1))
    grad = {k: _host(d) for k, d in solver.adjoint(da, a_up, a_dn).items()}
    ocl = O.Cloud(lut_lw)

    def loss(a):
        o = O.solve_lw(okd_lw, a, ocl)
        return float(np.sum(a_up * o["flux_up"]) + np.sum(a_dn * o["flux_dn"]))

    for trial in range(3):
        v = _directions(rng, atm, True)
        h = 1e-4
        fd = (loss(_perturbed(atm, v, h)) - loss(_perturbed(atm, v, -h))) / (2 * h)
        ad = _dot(G, grad, v, kd.gas_names, True)
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)


@pytest.mark.parametrize("top_at_1", [False, True])
def test_sw_adjoint_end_to_end_fd(G, O, syn, kd_sw, okd_sw, lut_sw, top_at_1):
    from conftest import AVAILABLE
    atm = syn.make_atmosphere(52, 8, 30, cloudy=True, top_at_1=top_at_1)
    kd = G.KDist(kd_sw, AVAILABLE, precision=8)
    solver = G.SolverSW(kd, G.Cloud(lut_sw, 8))
    da = G.DeviceAtmosphere(atm, 8)
    tsi = np.full(atm.ncol, atm.tsi / kd.scalar("tsi"))
    rng = np.random.default_rng(12)
    a_up, a_dn, a_dir = (rng.normal(size=(atm.ncol, atm.nlay + 1)) for _ in range(3))
    grad = {k: _host(d) for k, d in solver.adjoint(da, tsi, a_up, a_dn, a_dir).items()}
    ocl = O.Cloud(lut_sw)

    def loss(a):
        o = O.solve_sw(okd_sw, a, ocl, tsi_scaling=tsi)
        return float(np.sum(a_up * o["flux_up"]) + np.sum(a_dn * o["flux_dn"]) + np.sum(a_dir * o["flux_dn_dir"]))

    for trial in range(3):
        v = _directions(rng, atm, False)
        h = 1e-4
        fd = (loss(_perturbed(atm, v, h)) - loss(_perturbed(atm, v, -h))) / (2 * h)
        ad = _dot(G, grad, v, kd.gas_names, False)
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)


@pytest.mark.parametrize("prec", [8, 4])
def test_heating_rate_adjoint(G, prec):
    """rrtmgp_hip_heating_rate_adjoint_*: exact derivatives of hr = c (net_i - net_{i+1}) /
    (p_{i+1} - p_i), net = dn - up, against central differences of that formula in float64."""
    import ctypes as C
    from rrtmgp_amd import _lib
    from rrtmgp_amd.device import DeviceArray
    rng = np.random.default_rng(31)
    ncol, nlay = 5, 9
    dt = np.float64 if prec == 8 else np.float32
    up = rng.uniform(100, 400, (ncol, nlay + 1)).astype(dt)
    dn = rng.uniform(100, 400, (ncol, nlay + 1)).astype(dt)
    p = np.cumsum(rng.uniform(500, 5000, (ncol, nlay + 1)), axis=1).astype(dt)
    g = rng.normal(size=(ncol, nlay)).astype(dt)
    c = 9.80665 / 1004.64 * 86400.0

    def hr(u, d, pl):
        net = d.astype(np.float64) - u.astype(np.float64)
        return c * (net[:, :-1] - net[:, 1:]) / (pl[:, 1:].astype(np.float64) - pl[:, :-1])

    dev = [DeviceArray.from_host(x, dt) for x in (up, dn, p, g)]
    outs = [DeviceArray.from_host(np.full((ncol, nlay + 1), 0.5), dt) for _ in range(3)]   # the entry adds
    fn = getattr(_lib.hip(), "rrtmgp_hip_heating_rate_adjoint_" + ("f64" if prec == 8 else "f32"))
    assert fn(C.c_int64(ncol), C.c_int64(nlay), dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, outs[0].ptr,
              outs[1].ptr, outs[2].ptr, None) == 0
    gu, gd, gp = (o.to_host().astype(np.float64) - 0.5 for o in outs)
    up64, dn64, p64, g64 = (x.astype(np.float64) for x in (up, dn, p, g))
    tol = 1e-7 if prec == 8 else 2e-5          # FP64: the O(h^2) error of differencing 1/dp
    for name, base, got in (("up", up64, gu), ("dn", dn64, gd), ("p", p64, gp)):
        v = rng.normal(size=base.shape) * (10.0 if name == "p" else 1.0)
        h = 1e-3

        def f(s):
            args = {"up": up64, "dn": dn64, "p": p64}
            args[name] = base + s * v
            return float(np.sum(g64 * hr(args["up"], args["dn"], args["p"])))

        fd = (f(h) - f(-h)) / (2 * h)
        ad = float(np.sum(got * v))
        assert abs(ad - fd) <= tol * max(abs(fd), 1e-3), (name, ad, fd)

