"""torch.autograd path (rrtmgp_amd.torch_ops): gradients of a loss over LW and SW fluxes and
heating rates, back-propagated through the HIP adjoint kernels, against central differences of
the FP64 CPU oracle along random directions."""
import numpy as np
import pytest

from conftest import AVAILABLE
from test_gpu_adjoint import G, _directions, _perturbed   # noqa: F401  (G is a fixture)

pytestmark = pytest.mark.gpu


def _tensors(torch, atm, names, dev):
    return {k: torch.tensor(np.asarray(v, np.float64), device=dev, requires_grad=True) for k, v in names.items()}


@pytest.mark.parametrize("lw", [True, False])
def test_torch_fluxes_and_heating_rate_gradients(G, O, syn, kd_lw, kd_sw, okd_lw, okd_sw, lut_lw, lut_sw, lw):
    torch = pytest.importorskip("torch")
    from rrtmgp_amd import torch_ops
    dev = torch.device("cuda")
    atm = syn.make_atmosphere(61, 6, 24, cloudy=True)
    kd = G.KDist(kd_lw if lw else kd_sw, AVAILABLE, precision=8)
    cloud = G.Cloud(lut_lw if lw else lut_sw, 8)
    solver = G.SolverLW(kd, cloud) if lw else G.SolverSW(kd, cloud)
    clouds = {k: torch.tensor(getattr(atm, k), device=dev) for k in ("lwp", "iwp", "rel", "rei")}
    other = {k: v for k, v in atm.vmr.items() if k not in ("h2o", "o3")}
    x = _tensors(torch, atm, {"t_lay": atm.t_lay, "t_lev": atm.t_lev, "t_sfc": atm.t_sfc, "p_lay": atm.p_lay,
                              "p_lev": atm.p_lev, "h2o": atm.vmr["h2o"], "o3": atm.vmr["o3"],
                              "emis_sfc": atm.emis_sfc, "sfc_alb": atm.sfc_alb}, dev)
    rng = np.random.default_rng(21)
    nlev = atm.nlay + 1
    a = {k: rng.normal(size=(atm.ncol, nlev)) for k in ("up", "dn", "dir")}
    # Heating-rate cotangents only below 1000 Pa: above, layers of ~1 Pa make the heating rate a
    # quotient of near-cancelling flux differences, and the oracle's central differences of it
    # wander by ~1e-3 relative as the step shrinks (the adjoint agrees with its large-step limit).
    a_hr = rng.normal(size=(atm.ncol, atm.nlay)) * 10. * (atm.p_lay > 1000.)
    tsi = None if lw else np.full(atm.ncol, atm.tsi / kd.scalar("tsi"))
    if lw:
        up, dn = torch_ops.lw_fluxes(solver, x["t_lay"], x["t_lev"], x["t_sfc"], x["p_lay"], x["p_lev"], x["h2o"],
                                     x["o3"], x["emis_sfc"], other, clouds)
        extra = 0.
    else:
        up, dn, dr = torch_ops.sw_fluxes(solver, x["t_lay"], x["p_lay"], x["p_lev"], x["h2o"], x["o3"], x["sfc_alb"],
                                         torch.tensor(atm.mu0, device=dev), torch.tensor(tsi, device=dev), other,
                                         clouds)
        extra = (torch.tensor(a["dir"], device=dev) * dr).sum()
    hr = torch_ops.heating_rate(up, dn, x["p_lev"])
    T = lambda v: torch.tensor(v, device=dev)
    loss_t = (T(a["up"]) * up).sum() + (T(a["dn"]) * dn).sum() + (T(a_hr) * hr).sum() + extra
    loss_t.backward()
    grad = {k: v.grad.detach().cpu().numpy() if v.grad is not None else None for k, v in x.items()}

    ocl = O.Cloud(lut_lw if lw else lut_sw)
    okd = okd_lw if lw else okd_sw

    def loss(at):
        o = O.solve_lw(okd, at, ocl) if lw else O.solve_sw(okd, at, ocl, tsi_scaling=tsi)
        hr_ = O.heating_rate(o["flux_net"], at.p_lev)
        val = np.sum(a["up"] * o["flux_up"]) + np.sum(a["dn"] * o["flux_dn"]) + np.sum(a_hr * hr_)
        if not lw:
            val += np.sum(a["dir"] * o["flux_dn_dir"])
        return float(val)

    for trial in range(2):
        v = _directions(rng, atm, lw)
        # LW: the Planck table has a node every 1 K, so steps much above 1e-4 K cross kinks of its
        # linear interpolation; SW: at 1e-4 the FP64 two-stream's rounding at tiny tau shows.
        h = 1e-4 if lw else 1e-3
        fd = (loss(_perturbed(atm, v, h)) - loss(_perturbed(atm, v, -h))) / (2 * h)
        ad = (np.sum(grad["t_lay"] * v["t_lay"]) + np.sum(grad["p_lay"] * atm.p_lay * v["lnp_lay"]) +
              np.sum(grad["p_lev"] * v["p_lev"]) + np.sum(grad["h2o"] * v["h2o"]) + np.sum(grad["o3"] * v["o3"]))
        if lw:
            ad += (np.sum(grad["t_lev"] * v["t_lev"]) + np.sum(grad["t_sfc"] * v["t_sfc"]) +
                   np.sum(grad["emis_sfc"] * v["emis_sfc"]))
        else:
            ad += np.sum(grad["sfc_alb"] * v["sfc_alb"])
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)
