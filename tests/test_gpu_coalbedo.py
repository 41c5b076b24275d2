"""The shortwave chain in co-albedo form: gas optics, cloud optics (with and without delta-scaling),
aerosol optics, the band add_to and the band-summed solver carry 1 - ssa in place of ssa
(rrtmgp_hip_*_coalbedo_*, two_stream = 2, coalbedo = 1).  The shortwave driver runs this chain
whenever it has no surrogate, so every driver parity test exercises it too.

Checks:
  * FP64: each co-albedo output against 1 - ssa of the oracle's (ssa form, src/Gas_optics_rrtmgp.cpp,
    Cloud_optics.cpp, Aerosol_optics.cpp, Optical_props.cpp's inc_2stream_by_2stream_bybnd) within
    1e-13 absolute plus 1e-12 relative, tau and g within 1e-12 relative; solver fluxes within 1e-10.
  * FP32 gas optics: co-albedo within 1e-4 relative of the FP64 co-albedo on every element,
    including the Rayleigh-dominated ones where 1 - ssa is ~1e-8 and FP32 ssa cannot resolve it
    (its 1 - ssa is then off by its own size).  FP32 cloud optics: within 1e-4 relative of the
    oracle's 1 - ssa on the FP32-rounded table and inputs.
  * The FP32 chain's fluxes within 0.05 W m-2 of the oracle's and heating rates within 0.05 K/day.
"""
import numpy as np
import pytest

from conftest import AVAILABLE, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from rrtmgp_amd import api, device
    if device.device_count() < 1:
        pytest.fail("no HIP device visible")
    return api


def _h(d):
    return d.to_host().astype(np.float64)


def _close_co(got, want_ssa, what):
    want = 1.0 - want_ssa
    err = np.abs(got - want) - 1e-12 * np.abs(want)
    assert err.max() <= 1e-13, (what, float(err.max()))


def _sw_inputs(G, atm, prec):
    ncol, nlay = atm.p_lay.shape
    vmr = {g: np.broadcast_to(np.asarray(atm.vmr[g], np.float64), (ncol, nlay)).copy(order="F") for g in atm.vmr}
    return G.GasConcs(vmr, prec), G.get_col_dry(vmr["h2o"], atm.p_lev, prec)


def test_gas_optics_coalbedo(G, O, syn, kd_sw, okd_sw):
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 40, 128)
    print("gas optics co-albedo: 40 x 128", flush=True)
    ref = O.gas_optics(okd_sw, atm)
    co = {}
    for prec in (8, 4):
        kd = G.KDist(kd_sw, AVAILABLE, prec)
        gc, cd = _sw_inputs(G, atm, prec)
        o = G.gas_optics(kd, atm, gc, cd, coalbedo=True)
        co[prec] = _h(o["coalbedo"])
        if prec == 8:
            assert rel_err(_h(o["tau"]), ref["tau"], floor=1e-30) < 1e-12
            assert rel_err(_h(o["g"]), ref["g"], floor=1e-6) < 1e-12
            assert rel_err(_h(o["toa_src"]), ref["toa_src"], floor=1e-6) < 1e-12
            _close_co(co[8], ref["ssa"], "gas")
            plain = G.gas_optics(kd, atm, gc, cd)
            assert rel_err(_h(plain["tau"]), _h(o["tau"]), floor=1e-30) == 0.0
        else:
            f32 = G.gas_optics(kd, atm, gc, cd)
            ssa32 = _h(f32["ssa"])
    small = co[8] < 1e-6
    assert small.any(), "no nearly conservative gas layers in the case"
    assert rel_err(co[4], co[8], floor=1e-30) < 1e-4
    # What the ssa form gives in FP32 where 1 - ssa is small: its own size off.
    worst = np.abs((1.0 - ssa32[small]) - co[8][small]) / co[8][small]
    print(f"FP32 1-ssa relative error where 1-ssa < 1e-6: max {worst.max():.2e}; "
          f"co-albedo form {rel_err(co[4][small], co[8][small], floor=1e-30):.2e}", flush=True)


def _round32(lut):
    import dataclasses
    return dataclasses.replace(lut, **{f.name: np.asarray(getattr(lut, f.name), np.float32).astype(np.float64)
                                       for f in dataclasses.fields(lut)
                                       if isinstance(getattr(lut, f.name), np.ndarray)})


@pytest.mark.parametrize("delta", [False, True])
def test_cloud_optics_coalbedo(G, O, syn, lut_sw, delta):
    """FP64 against the oracle; FP32 against the oracle on the FP32-rounded table and inputs (the
    table's ssa is an FP32 input there, so 1 - ssa can be no better than 1 - its rounding)."""
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 33, 60, cloudy=True)
    print(f"cloud optics co-albedo: delta={delta}", flush=True)
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)    # noqa: E731
    for prec in (8, 4):
        lut, cl = (lut_sw, (atm.lwp, atm.iwp, atm.rel, atm.rei)) if prec == 8 else \
            (_round32(lut_sw), tuple(r32(a) for a in (atm.lwp, atm.iwp, atm.rel, atm.rei)))
        t, s, g = O.Cloud(lut).optics(*cl, True)
        if delta:
            t, s, g = O.delta_scale(t, s, g)
        c = G.Cloud(lut_sw, prec)
        gt, gc, gg = (_h(a) for a in c.optics(*cl, True, delta_scale=delta, coalbedo=True))
        if prec == 8:
            assert rel_err(gt, t, floor=1e-30) < 1e-12
            assert rel_err(gg, g, floor=1e-6) < 1e-12
            _close_co(gc, s, "cloud")
            # Clear layers: tau = 0, co-albedo 1 (ssa 0), as the ssa form writes ssa = 0 there.
            assert (gc[t == 0] == 1.0).all()
        else:
            m = t > 0
            assert m.sum() > 100
            assert rel_err(gc[m], 1.0 - s[m], floor=1e-30) < 1e-4
            assert rel_err(gt, t, floor=1e-30) < 1e-5


@pytest.mark.parametrize("delta", [False, True])
def test_aerosol_optics_coalbedo(G, O, syn, delta):
    import os
    from conftest import ROOT
    from rrtmgp_amd import coefficients, netcdf
    lut = coefficients.aerosol_lut(netcdf.Dataset(os.path.join(ROOT, "tests", "golden", "aerosol_optics.nc")))
    atm = syn.make_atmosphere(61, 21, 30)
    ncol, nlay = atm.p_lay.shape
    rng = np.random.default_rng(8)
    f = {f"aermr{i:02d}": rng.uniform(1e-10, 1e-7, (ncol, nlay)) for i in range(1, 12)}
    rh = rng.uniform(0.0, 1.0, (ncol, nlay))
    t, s, g = O.aerosol_optics(f, rh, atm.p_lev, lut)
    if delta:
        t, s, g = O.delta_scale(t, s, g)
    a = G.Aerosol(lut, precision=8)
    gt, gc, gg = (_h(x) for x in a.optics(G.GasConcs(f, precision=8), rh, atm.p_lev, delta_scale=delta,
                                          coalbedo=True))
    assert rel_err(gt, t, floor=1e-30) < 1e-12
    assert rel_err(gg, g, floor=1e-6) < 1e-12
    _close_co(gc, s, "aerosol")


@pytest.mark.parametrize("prec", [8, 4])
def test_add_to_and_solver_coalbedo(G, O, syn, kd_sw, okd_sw, lut_sw, prec):
    """Gas + delta-scaled cloud added in co-albedo form, then the band-summed solver, against the
    oracle's ssa-form chain; FP32 within 0.05 W m-2 and its heating rates within 0.05 K/day of the
    FP64 co-albedo chain's."""
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 36, 128, cloudy=True)
    print(f"co-albedo add_to + solver: prec={prec}", flush=True)
    kd = G.KDist(kd_sw, AVAILABLE, prec)
    gc, cd = _sw_inputs(G, atm, prec)
    go = G.gas_optics(kd, atm, gc, cd, coalbedo=True)
    ct, cc, cg = G.Cloud(lut_sw, prec).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True, delta_scale=True,
                                             coalbedo=True)
    G.add_to(kd, go["tau"], ct, go["coalbedo"], go["g"], cc, cg, coalbedo=True)
    ref = O.gas_optics(okd_sw, atm)
    t, s, g = O.delta_scale(*O.Cloud(lut_sw).optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    bl = np.asarray(okd_sw.band_lims)
    tau, ssa, gg = O.inc_bybnd(ref["tau"], t, bl, ref["ssa"], ref["g"], s, g)
    if prec == 8:
        assert rel_err(_h(go["tau"]), tau, floor=1e-30) < 1e-12
        assert rel_err(_h(go["g"]), gg, floor=1e-6) < 1e-12
        _close_co(_h(go["coalbedo"]), ssa, "add_to")
    scale = atm.tsi / kd.scalar("tsi")
    from rrtmgp_amd import device
    toa = device.DeviceArray.from_host(go["toa_src"].to_host() * scale, G.real_dtype(prec))
    o = G.rte_sw_bnd(kd, atm.top_at_1, go["tau"], go["coalbedo"], go["g"], atm.mu0, toa, atm.sfc_alb,
                     0.5 * atm.sfc_alb, p_lev=atm.p_lev, coalbedo=True)
    up, dn, dr = O.rte_sw(tau, ssa, gg, atm.mu0, ref["toa_src"] * scale, atm.sfc_alb, 0.5 * atm.sfc_alb, bl,
                          atm.top_at_1, do_broadband=True)
    for name, want in (("flux_up", up), ("flux_dn", dn), ("flux_dn_dir", dr)):
        got = _h(o[name])
        if prec == 8:
            assert rel_err(got, want, floor=1e-6) < 1e-10, name
        else:
            assert np.abs(got - want).max() < 0.05, name
    hr = O.heating_rate(dn - up, atm.p_lev)
    d = np.abs(_h(o["heating_rate"]) - hr)
    if prec == 8:
        fmax = max(np.abs(up).max(), np.abs(dn).max())
        dp = np.abs(np.diff(atm.p_lev, axis=1))
        assert np.all(d <= 1e-8 * np.abs(hr) + 9.80665 / 1004.64 * 86400.0 * 2e-10 * fmax / dp)
    else:
        print(f"FP32 co-albedo heating rate: max |err| {d.max():.3e} K/day", flush=True)
        assert d.max() < 0.05
