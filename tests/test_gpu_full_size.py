"""The all-sky LW+SW step at its full size, 65536 columns x 128 layers in FP32: the g-point arrays
hold 65536 * 128 * 256 = 2^31 elements (8 GiB), so the offsets of the last columns' g-points pass
2^31 elements and 4 GiB.  Checked by two properties that do not depend on the size:

* 128 columns (the last 64 and 64 spread over the rest) give the same fluxes, bit for bit, as the
  same columns run on their own (every kernel of the step is column-independent, so a wrong offset
  at large indices shows up as a difference);
* those columns agree with the FP64 oracle within the FP32 parity bar, 0.05 W m-2.

The runs go through the class API's drivers (Radiation_solver_{longwave,shortwave}::solve_gpu) as
the bench does, on the same synthetic workload.
"""
import numpy as np
import pytest

from conftest import AVAILABLE

pytestmark = pytest.mark.gpu

NCOL, NLAY, PREC = 65536, 128, 4


def _step(api, device, atm, kd_lw, kd_sw, lut_lw, lut_sw):
    ncol, nlev = atm.ncol, atm.nlay + 1
    lw = api.SolverLW(kd_lw, api.Cloud(lut_lw, PREC))
    sw = api.SolverSW(kd_sw, api.Cloud(lut_sw, PREC))
    da = api.DeviceAtmosphere(atm, PREC)
    dt = api.real_dtype(PREC)
    tsi = device.DeviceArray.from_host(np.full(ncol, atm.tsi / kd_sw.scalar("tsi")), dt)
    out_lw = {k: device.DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_net")}
    out_sw = {k: device.DeviceArray((ncol, nlev), dt) for k in ("flux_up", "flux_dn", "flux_dn_dir", "flux_net")}
    lw.solve(da, outputs=out_lw)
    sw.solve(da, tsi, outputs=out_sw)
    device.synchronize()
    return ({"lw_" + k: v.to_host() for k, v in out_lw.items()} |
            {"sw_" + k: v.to_host() for k, v in out_sw.items()})


def test_full_size_step_columns(syn, kd_lw, kd_sw, lut_lw, lut_sw, O):
    from rrtmgp_amd import api, device
    if device.device_count() < 1:
        pytest.fail("no HIP device visible")
    klw, ksw = api.KDist(kd_lw, AVAILABLE, PREC), api.KDist(kd_sw, AVAILABLE, PREC)
    atm = syn.make_atmosphere(syn.SEED_ATMOS[1], NCOL, NLAY, cloudy=True)
    print("full-size step", NCOL, NLAY, flush=True)
    full = _step(api, device, atm, klw, ksw, lut_lw, lut_sw)
    cols = np.concatenate([np.linspace(0, NCOL - 65, 64).astype(int), np.arange(NCOL - 64, NCOL)])
    sub = syn.subset_columns(atm, cols)
    print("the same columns alone", sub.ncol, flush=True)
    part = _step(api, device, sub, klw, ksw, lut_lw, lut_sw)
    for k in full:
        assert np.all(np.isfinite(full[k])), k
        assert np.array_equal(full[k][cols], part[k]), k
    okl, oks = O.Kdist(kd_lw, AVAILABLE), O.Kdist(kd_sw, AVAILABLE)
    rlw = O.solve_lw(okl, sub, O.Cloud(lut_lw))
    rsw = O.solve_sw(oks, sub, O.Cloud(lut_sw), tsi_scaling=np.full(sub.ncol, sub.tsi / oks.scalar("tsi")))
    for k in ("flux_up", "flux_dn"):
        assert np.abs(part["lw_" + k] - rlw[k]).max() < 0.05, "lw_" + k
    for k in ("flux_up", "flux_dn", "flux_dn_dir"):
        assert np.abs(part["sw_" + k] - rsw[k]).max() < 0.05, "sw_" + k
