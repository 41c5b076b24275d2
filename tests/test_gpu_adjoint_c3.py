"""The adjoint at its own workload: all-sky inputs, 16384 columns x 128 layers, FP32 forward and
FP32 end-to-end adjoint (rrtmgp_api_adjoint_lw / _sw: per-band gas-optics recomputation, solver
adjoint, gas-optics adjoint, col_dry chain).

Cotangents of the broadband up and down (and SW direct) fluxes are random on 256 random columns
and zero elsewhere; since columns are independent, the gradient lives on those columns only.
Checked there, for every input group (T_lay, T_lev, T_sfc, ln p_lay, p_lev, h2o and o3 vmr,
emissivity / albedo):
  * against the FP64 adjoint of the same columns, within 1e-3 of the group's largest gradient;
  * against central differences of the FP64 oracle (step 1e-4 relative to each input) along
    random directions, per input group and for all groups at once, within 1e-3 of the sum of
    the direction's absolute contributions sum |g v|.
The interpolation switches tables at the tropopause pressure, a jump in ln p; the 256 columns are
drawn from those with no layer within 10 steps of it, so no difference crosses the jump.  The
gas optics is piecewise linear in temperature (1 K Planck-table nodes, which include the 15 K
k-table nodes) and in ln p (k-table pressure nodes): a central difference that straddles a node
averages two slopes, which no derivative reproduces.  Unmasked, such crossings leave the FP64
oracle's own central differences unconverged at 1e-4 (3e-3 between steps 1e-4 and 1e-5 for T_lay,
1.5e-3 for T_lev); so each direction leaves out the values within six standard deviations of
the step of a node, after which the oracle's differences agree between steps 1e-4, 1e-5 and
1e-6 to 2e-7.
"""
import copy

import numpy as np
import pytest

from conftest import AVAILABLE

pytestmark = pytest.mark.gpu

NCOL, NLAY, NSAMPLE, H = 16384, 128, 256, 1e-4


@pytest.fixture(scope="module")
def c3(syn, kd_lw):
    atm = syn.make_atmosphere(syn.SEED_ATMOS[3], NCOL, NLAY, cloudy=True)
    rng = np.random.default_rng(33)
    lnp_trop = np.log(kd_lw.press_ref_trop)
    safe = np.all(np.abs(np.log(atm.p_lay) - lnp_trop) > 10 * H * 5, axis=1)
    cols = np.sort(rng.choice(np.flatnonzero(safe), NSAMPLE, replace=False))
    return atm, cols, rng


def _clear_of_nodes(x, x0, dx, step):
    """True where x lies more than six standard deviations of the step from the nodes x0 + n dx."""
    f = np.mod((x - x0) / dx, 1.0)
    return np.minimum(f, 1.0 - f) * abs(dx) > 6.0 * np.abs(step)


def _relative_directions(rng, a, lw, kd):
    """Perturbation directions scaled to the inputs (x -> x + h r x, r ~ N(0, 1)); pressures at the
    table's lower edge (1 Pa) are held, since the oracle rejects values below it, and temperatures
    and layer pressures near an interpolation node are held (see the module notes)."""
    t0 = float(np.min(kd.temp_ref))
    dT = (float(np.max(kd.temp_ref)) - t0) / (kd.totplnk.shape[0] - 1) if kd.is_lw else \
        float(kd.temp_ref[1] - kd.temp_ref[0])
    lnp_ref = np.log(np.asarray(kd.press_ref))
    dlnp = (lnp_ref[-1] - lnp_ref[0]) / (len(lnp_ref) - 1)
    clear_t = lambda t: _clear_of_nodes(t, t0, dT, H * t)      # noqa: E731
    v = {"t_lay": rng.normal(size=a.t_lay.shape) * a.t_lay * clear_t(a.t_lay),
         "lnp_lay": rng.normal(size=a.p_lay.shape) * (a.p_lay > 2.) *
                    _clear_of_nodes(np.log(a.p_lay), lnp_ref[0], dlnp, H),
         "p_lev": rng.normal(size=a.p_lev.shape) * a.p_lev * (a.p_lev > 2.),
         "h2o": rng.normal(size=a.vmr["h2o"].shape) * a.vmr["h2o"],
         "o3": rng.normal(size=a.vmr["o3"].shape) * a.vmr["o3"]}
    if lw:
        v.update({"t_lev": rng.normal(size=a.t_lev.shape) * a.t_lev * clear_t(a.t_lev),
                  "t_sfc": rng.normal(size=a.t_sfc.shape) * a.t_sfc * clear_t(a.t_sfc) * clear_t(a.t_sfc + 1.),
                  "emis_sfc": rng.normal(size=a.emis_sfc.shape) * 0.01})
    else:
        v["sfc_alb"] = rng.normal(size=a.sfc_alb.shape) * a.sfc_alb
    return v


def _perturbed(atm, v, h):
    a = copy.deepcopy(atm)
    a.t_lay = atm.t_lay + h * v["t_lay"]
    a.p_lay = atm.p_lay * np.exp(h * v["lnp_lay"])
    a.p_lev = atm.p_lev + h * v["p_lev"]
    a.vmr = dict(atm.vmr)
    for gas in ("h2o", "o3"):
        a.vmr[gas] = atm.vmr[gas] + h * v[gas]
    if "t_lev" in v:
        a.t_lev, a.t_sfc, a.emis_sfc = atm.t_lev + h * v["t_lev"], atm.t_sfc + h * v["t_sfc"], \
            atm.emis_sfc + h * v["emis_sfc"]
    if "sfc_alb" in v:
        a.sfc_alb = atm.sfc_alb + h * v["sfc_alb"]
    return a


def _groups(g, names, lw, cols=None):
    sel = (lambda x: x) if cols is None else (lambda x: x[cols])
    out = {k: sel(g[k]) for k in ("t_lay", "lnp_lay", "p_lev")}
    out["h2o"] = sel(g["vmr"][:, :, names.index("h2o")])
    out["o3"] = sel(g["vmr"][:, :, names.index("o3")])
    if lw:
        out.update({"t_lev": sel(g["t_lev"]), "t_sfc": sel(g["t_sfc"]),
                    "emis_sfc": g["emis_sfc"][:, cols] if cols is not None else g["emis_sfc"]})
    else:
        ad, af = g["sfc_alb_dir"], g["sfc_alb_dif"]
        out["sfc_alb"] = (ad + af)[:, cols] if cols is not None else ad + af
    return out


@pytest.mark.parametrize("lw", [True, False])
def test_adjoint_c3_workload(c3, syn, O, okd_lw, okd_sw, kd_lw, kd_sw, lut_lw, lut_sw, lw):
    from rrtmgp_amd import api
    atm, cols, rng = c3
    sub = syn.subset_columns(atm, cols)
    nlev = NLAY + 1
    ncot = 2 if lw else 3
    cot_sub = [rng.normal(size=(NSAMPLE, nlev)) for _ in range(ncot)]
    cot_full = []
    for c in cot_sub:
        full = np.zeros((NCOL, nlev))
        full[cols] = c
        cot_full.append(np.asfortranarray(full))
    print(f"C3 adjoint {'LW' if lw else 'SW'}: FP32 at {NCOL} x {NLAY}, {NSAMPLE} sampled columns", flush=True)

    grads = {}
    for prec, a, cot in ((4, atm, cot_full), (8, sub, [np.asfortranarray(c) for c in cot_sub])):
        kd = api.KDist(kd_lw if lw else kd_sw, AVAILABLE, prec)
        da = api.DeviceAtmosphere(a, prec)
        if lw:
            g = api.SolverLW(kd, api.Cloud(lut_lw, prec)).adjoint(da, *cot)
        else:
            tsi = np.full(a.ncol, a.tsi / kd.scalar("tsi"))
            g = api.SolverSW(kd, api.Cloud(lut_sw, prec)).adjoint(da, tsi, *cot)
        g = {k: d.to_host().astype(np.float64) for k, d in g.items()}
        grads[prec] = _groups(g, kd.gas_names, lw, cols if prec == 4 else None)
        names = kd.gas_names
        del da

    # FP32 against FP64 adjoint, per input group.
    for k in grads[8]:
        g32, g64 = grads[4][k], grads[8][k]
        scale = np.abs(g64).max()
        err = np.abs(g32 - g64).max()
        print(f"  {k}: max |g32 - g64| / max |g64| = {err / scale:.2e}", flush=True)
        assert err <= 1e-3 * scale, (k, err, scale)

    # FP32 adjoint against central differences of the FP64 oracle on the sampled columns.
    if lw:
        ocl = O.Cloud(lut_lw)

        def loss(a):
            o = O.solve_lw(okd_lw, a, ocl)
            return float(np.sum(cot_sub[0] * o["flux_up"]) + np.sum(cot_sub[1] * o["flux_dn"]))
    else:
        ocl = O.Cloud(lut_sw)
        tsi = np.full(NSAMPLE, sub.tsi / okd_sw.scalar("tsi"))

        def loss(a):
            o = O.solve_sw(okd_sw, a, ocl, tsi_scaling=tsi)
            return float(np.sum(cot_sub[0] * o["flux_up"]) + np.sum(cot_sub[1] * o["flux_dn"]) +
                         np.sum(cot_sub[2] * o["flux_dn_dir"]))

    # Each input group alone, then all groups together.  The error is measured against the sum
    # of the absolute contributions |g v| of the direction: a random direction's derivative can
    # cancel to a small fraction of its terms, and the FD error (the eta nodes, which the masks
    # do not remove) scales with the terms, not with their sum.
    g = grads[4]
    v_all = _relative_directions(rng, sub, lw, kd_lw if lw else kd_sw)
    trials = [(k, {j: (v_all[j] if j == k else 0 * v_all[j]) for j in v_all}) for k in v_all]
    trials.append(("all", v_all))
    for name, v in trials:
        scale = sum(float(np.sum(np.abs(g[k] * v[k]))) for k in v)
        assert scale > 0, name
        fd = (loss(_perturbed(sub, v, H)) - loss(_perturbed(sub, v, -H))) / (2 * H)
        ad = sum(float(np.sum(g[k] * v[k])) for k in v)
        print(f"  direction {name}: adjoint {ad:.8e}, central difference {fd:.8e}, "
              f"|ad - fd| / sum|g v| = {abs(ad - fd) / scale:.2e}", flush=True)
        assert abs(ad - fd) <= 1e-3 * scale, (name, ad, fd, scale)
