"""Pins the oracle's gas optics (init, interpolation, absorption, Rayleigh, Planck) with closed
forms: tables chosen so the interpolation collapses to a known function of the inputs."""
import dataclasses

import numpy as np
import pytest

from conftest import AVAILABLE, rel_err


def _flat_vmr_ref(kd):
    """vmr_ref independent of temperature, so col_mix is the same for both temperature nodes."""
    v = kd.vmr_ref.copy(order="F")
    v[:, :, :] = v[:, :, :1]
    return v


def _variant(kd, kmajor=None, kminor=None, krayl=None, flat=True):
    ch = {}
    if flat:
        ch["vmr_ref"] = _flat_vmr_ref(kd)
    if kmajor is not None:
        ch["kmajor"] = np.asfortranarray(np.broadcast_to(kmajor, kd.kmajor.shape).astype(np.float64))
    if kminor is not None:
        ch["kminor_lower"] = np.asfortranarray(np.broadcast_to(kminor, kd.kminor_lower.shape).astype(np.float64))
        ch["kminor_upper"] = np.asfortranarray(np.broadcast_to(kminor, kd.kminor_upper.shape).astype(np.float64))
    if krayl is not None:
        ch["rayl_lower"] = np.asfortranarray(np.broadcast_to(krayl, kd.rayl_lower.shape).astype(np.float64))
        ch["rayl_upper"] = np.asfortranarray(np.broadcast_to(krayl, kd.rayl_upper.shape).astype(np.float64))
    return dataclasses.replace(kd, **ch)


@pytest.fixture(scope="module")
def col_dry(O, atm_c0):
    return O.get_col_dry(atm_c0.vmr["h2o"], atm_c0.p_lev)


def test_col_dry_integrates_to_surface_pressure(O, atm_c0, col_dry):
    # sum_k col_dry*(1+h2o)*m_air*g/N_A  (molec/cm2 -> kg/m2 * g) = p_sfc - p_top  (Pa)
    h2o = atm_c0.vmr["h2o"]
    m_air = (0.028964 + 0.018016 * h2o) / (1 + h2o)
    mass = col_dry * (1 + h2o) * m_air * 1e4 / 6.02214076e23 * 9.80665
    assert rel_err(mass.sum(1), atm_c0.p_lev[:, 0] - atm_c0.p_lev[:, -1]) < 1e-12


def test_flavors_by_hand(okd_lw, okd_sw):
    # Unique (gas1, gas2) pairs of LW_KEY_LOWER/UPPER in band order, (0,0) -> (2,2).
    lw = [(1, 0), (1, 2), (2, 3), (1, 3), (3, 2), (3, 0), (1, 4), (4, 2), (2, 2)]
    sw = [(1, 2), (2, 0), (1, 0), (1, 3), (3, 2), (3, 0), (2, 2)]
    assert okd_lw.get_int("flavor").reshape(-1, 2).tolist() == [list(p) for p in lw]
    assert okd_sw.get_int("flavor").reshape(-1, 2).tolist() == [list(p) for p in sw]
    gf = okd_lw.get_int("gpoint_flavor").reshape(-1, 2)
    assert gf.shape == (256, 2)
    assert (gf[0:16] == [1, 1]).all()           # band 1: (h2o,0) lower and upper
    assert (gf[240:256] == [2, 9]).all()        # band 16: (h2o,co2) lower, (0,0)->(2,2) upper
    assert (gf[32:48] == [2, 3]).all()          # band 3: (h2o,co2), (co2,o3)


def test_minor_reduction_drops_missing_gas(O, kd_lw):
    full = O.Kdist(kd_lw, AVAILABLE)
    red = O.Kdist(kd_lw, [g for g in AVAILABLE if g != "ch4"])
    keep = [i for i, m in enumerate(kd_lw.minor_gases_lower) if m != "ch4"]
    assert red.nminor_lower == len(keep) < full.nminor_lower
    assert red.gas_names == ["h2o", "co2", "o3", "n2o"]
    lim = red.get_int("minor_limits_gpt_lower").reshape(-1, 2)
    assert (lim == kd_lw.minor_limits_gpt_lower.T[keep]).all()
    start = red.get_int("kminor_start_lower")
    widths = lim[:, 1] - lim[:, 0] + 1
    assert (start == np.concatenate([[1], 1 + np.cumsum(widths)[:-1]])).all()
    # Kept contributor rows are copied unchanged (tables reordered to (ntemp, neta, ncontrib)).
    kf = full.get_real("kminor_lower").reshape(red.ncontrib_lower * 0 + full.ncontrib_lower, 9, 14)
    kr = red.get_real("kminor_lower").reshape(red.ncontrib_lower, 9, 14)
    src = kd_lw.kminor_lower   # (ncontrib, neta, ntemp)
    for j, i in enumerate(keep):
        s0 = kd_lw.kminor_start_lower[i] - 1
        r0 = start[j] - 1
        np.testing.assert_array_equal(kr[r0:r0 + 16], src[s0:s0 + 16].transpose(0, 1, 2))
    assert kf.shape[0] == kd_lw.kminor_lower.shape[0]


def test_missing_key_gas_raises(O, kd_lw):
    with pytest.raises(RuntimeError, match="required gas co2 is missing"):
        O.Kdist(kd_lw, ["h2o", "o3", "n2o", "ch4"])


def test_interpolation_invariants(O, okd_lw, atm_c0, col_dry):
    o = O.gas_optics(okd_lw, atm_c0, col_dry=col_dry, interp=True)
    assert o["jtemp"].min() >= 1 and o["jtemp"].max() <= okd_lw.ntemp - 1
    assert o["jpress"].min() >= 1 and o["jpress"].max() <= okd_lw.npres - 1
    assert o["jeta"].min() >= 1 and o["jeta"].max() <= okd_lw.neta - 1
    np.testing.assert_allclose(o["fmajor"].sum(axis=(0, 1, 2)), 1.0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(o["fminor"].sum(axis=(0, 1)), 1.0, rtol=0, atol=1e-14)
    tropo = np.log(atm_c0.p_lay) > okd_lw.scalar("press_ref_trop_log")
    assert (o["tropo"].astype(bool) == tropo).all()


def _tau(O, kd, atm, col_dry, avail=AVAILABLE):
    return O.gas_optics(O.Kdist(kd, avail), atm, col_dry=col_dry)["tau"]


def test_major_linear_in_temperature(O, kd_lw, atm_c0, col_dry):
    # k(t) = T_ref(t): the two-point temperature interpolation returns T_lay exactly.
    kt = kd_lw.temp_ref[None, None, None, :]
    t1 = _tau(O, _variant(kd_lw, kmajor=1.0, kminor=0.0), atm_c0, col_dry)
    tt = _tau(O, _variant(kd_lw, kmajor=kt, kminor=0.0), atm_c0, col_dry)
    assert rel_err(tt / t1, np.broadcast_to(atm_c0.t_lay[:, :, None], tt.shape)) < 1e-13


def test_major_linear_in_log_pressure(O, kd_lw, atm_c0, col_dry, okd_lw):
    # k(p) = p (0-based node index): interpolation returns the fractional pressure index,
    # shifted by one node in the upper atmosphere (the tropo offset of jpress).
    npres = kd_lw.press_ref.size
    kp = np.arange(npres + 1, dtype=np.float64)[None, None, :, None]
    t1 = _tau(O, _variant(kd_lw, kmajor=1.0, kminor=0.0), atm_c0, col_dry)
    tp = _tau(O, _variant(kd_lw, kmajor=kp, kminor=0.0), atm_c0, col_dry)
    lnp = np.log(atm_c0.p_lay)
    d = (np.log(kd_lw.press_ref[-1]) - np.log(kd_lw.press_ref[0])) / (npres - 1)
    upper = lnp <= np.log(kd_lw.press_ref_trop)
    expect = (lnp - np.log(kd_lw.press_ref[0])) / d + upper
    assert rel_err(tp / t1, np.broadcast_to(expect[:, :, None], tp.shape), floor=1.0) < 1e-12


def test_major_linear_in_eta(O, kd_lw, atm_c0, col_dry, okd_lw):
    # k(eta) = eta: sum_i col_mix_i * ftemp_i * eta_i = column of the first key gas, for any T.
    neta = kd_lw.kmajor.shape[1]
    ke = (np.arange(neta) / (neta - 1.0))[None, :, None, None]
    te = _tau(O, _variant(kd_lw, kmajor=ke, kminor=0.0, flat=False), atm_c0, col_dry)
    flav = okd_lw.get_int("flavor").reshape(-1, 2)
    gf = okd_lw.get_int("gpoint_flavor").reshape(-1, 2)
    names = okd_lw.gas_names
    upper = np.log(atm_c0.p_lay) <= np.log(kd_lw.press_ref_trop)
    expect = np.zeros_like(te)
    for g in range(te.shape[2]):
        for half in (0, 1):
            gas1 = flav[gf[g, half] - 1, 0]
            col1 = np.broadcast_to(atm_c0.vmr[names[gas1 - 1]], col_dry.shape) * col_dry
            m = upper if half == 1 else ~upper
            expect[:, :, g][m] = col1[m]
    assert rel_err(te, expect) < 1e-13


def test_minor_constant_table_gives_scaling(O, kd_lw, atm_c0, col_dry, okd_lw):
    # kmajor = 0, kminor = 1: tau(g) = sum of the scalings of the active minor intervals.
    tm = _tau(O, _variant(kd_lw, kmajor=0.0, kminor=1.0), atm_c0, col_dry)
    t2 = _tau(O, _variant(kd_lw, kmajor=0.0, kminor=2.0), atm_c0, col_dry)
    assert rel_err(t2, 2 * tm, floor=1e-300) < 1e-14     # linear in kminor
    vmr = {k: np.broadcast_to(v, col_dry.shape) for k, v in atm_c0.vmr.items()}
    upper = np.log(atm_c0.p_lay) <= np.log(kd_lw.press_ref_trop)
    gas_of = dict(zip(kd_lw.identifier_minor, kd_lw.gas_minor))
    expect = np.zeros_like(tm)
    for side, mask in (("lower", ~upper), ("upper", upper)):
        for i, ident in enumerate(getattr(kd_lw, f"minor_gases_{side}")):
            s = vmr[gas_of[ident]] * col_dry
            if getattr(kd_lw, f"minor_scales_with_density_{side}")[i]:
                s = s * atm_c0.p_lay / 100.0 / atm_c0.t_lay
                sg = getattr(kd_lw, f"scaling_gas_{side}")[i]
                if sg:
                    frac = vmr[sg] / (1.0 + vmr["h2o"])
                    s = s * ((1.0 - frac) if getattr(kd_lw, f"scale_by_complement_{side}")[i] else frac)
            g0, g1 = getattr(kd_lw, f"minor_limits_gpt_{side}")[:, i]
            expect[:, :, g0 - 1:g1] += np.where(mask, s, 0.0)[:, :, None]
    assert rel_err(tm, expect, floor=1e-300) < 1e-13


def test_rayleigh_closed_form(O, kd_sw, atm_c0, col_dry):
    # No absorption and krayl(t) = T_ref(t) * c: tau = c*T_lay*(col_h2o + col_dry), ssa = 1, g = 0.
    c = 1e-27
    kr = c * kd_sw.temp_ref[None, None, :]
    okd = O.Kdist(_variant(kd_sw, kmajor=0.0, kminor=0.0, krayl=kr), AVAILABLE)
    o = O.gas_optics(okd, atm_c0, col_dry=col_dry)
    col_h2o = atm_c0.vmr["h2o"] * col_dry
    expect = (c * atm_c0.t_lay * (col_h2o + col_dry))[:, :, None]
    assert rel_err(o["tau"], np.broadcast_to(expect, o["tau"].shape)) < 1e-13
    assert np.all(o["ssa"] == 1.0) and np.all(o["g"] == 0.0)
    np.testing.assert_allclose(o["toa_src"], np.broadcast_to(okd.get_real("solar_source"), o["toa_src"].shape))


def test_combine_ssa_is_rayleigh_fraction(O, okd_sw, kd_sw, atm_c0, col_dry):
    o = O.gas_optics(okd_sw, atm_c0, col_dry=col_dry)
    oa = O.gas_optics(O.Kdist(_variant(kd_sw, krayl=0.0, flat=False), AVAILABLE), atm_c0, col_dry=col_dry)
    tau_ray = o["tau"] - oa["tau"]     # absolute error ~eps*tau, so compare ssa absolutely
    assert np.max(np.abs(o["ssa"] - tau_ray / o["tau"])) < 1e-13
    assert o["ssa"].min() >= 0.0 and o["ssa"].max() <= 1.0


def test_planck_band_sum_rule(O, okd_lw, kd_lw, atm_c0, col_dry):
    # planck_frac sums to one over each band's g-points at every node, so the band sum of each
    # source is the linearly interpolated band Planck function.
    o = O.gas_optics(okd_lw, atm_c0, col_dry=col_dry)
    tgrid = np.linspace(160.0, 355.0, kd_lw.totplnk.shape[0])
    b2g = kd_lw.band2gpt
    for b in range(kd_lw.nbnd):
        sl = slice(b2g[0, b] - 1, b2g[1, b])
        B = lambda t: np.interp(t, tgrid, kd_lw.totplnk[:, b])
        assert rel_err(o["lay_src"][:, :, sl].sum(2), B(atm_c0.t_lay)) < 1e-12
        assert rel_err(o["lev_src_inc"][:, :, sl].sum(2), B(atm_c0.t_lev[:, 1:])) < 1e-12
        assert rel_err(o["lev_src_dec"][:, :, sl].sum(2), B(atm_c0.t_lev[:, :-1])) < 1e-12
        assert rel_err(o["sfc_src"][:, sl].sum(1), B(atm_c0.t_sfc)) < 1e-12
        assert rel_err(o["sfc_src_jac"][:, sl].sum(1), B(atm_c0.t_sfc + 1) - B(atm_c0.t_sfc)) < 1e-9
    # pi * sum over bands of the band Planck radiance is sigma*T^4 (Stefan-Boltzmann).
    sigma = 5.670374419e-8
    tot = np.pi * o["lay_src"].sum(2)
    assert rel_err(tot, sigma * atm_c0.t_lay ** 4) < 2e-3


def test_range_checks(O, okd_lw, atm_c0, syn):
    bad = syn.make_atmosphere(7, 4, 20)
    bad.t_lay[0, 3] = 100.0
    with pytest.raises(ValueError, match="tlay is out of range"):
        O.gas_optics(okd_lw, bad)
    bad = syn.make_atmosphere(7, 4, 20)
    bad.p_lay[1, 2] = 2e5
    with pytest.raises(ValueError, match="play is out of range"):
        O.gas_optics(okd_lw, bad)
    bad = syn.make_atmosphere(7, 4, 20)
    cd = -np.ones((4, 20))
    with pytest.raises(ValueError, match="col_dry is out of range"):
        O.gas_optics(okd_lw, bad, col_dry=cd)


# ---------------------------------------------------------------------------------------------
# Closed forms of the three table gathers that sum rules leave open: the Planck fraction
# (src_kernels_cuda/gas_optics_rrtmgp_kernels.cu:181-190), k-minor (:451-455) and Rayleigh
# (:572-616).  Each table is c(index) * (1 + a T_ref(t)/300 + b eta(e) + d p(node)), with c
# varying over g-points or contributors and the slopes differing between the lower and upper
# tables.  Multilinear interpolation reproduces a function linear in each node coordinate, so the
# result is c * (1 + a T/300 + b eta_T + d locp) with eta_T = (1-ftemp) eta_0 + ftemp eta_1
# (eta_i at temperature node jtemp+i).  A swapped corner, a j0/j1 mix-up, a lost jpress+itropo
# shift, the wrong flavor or contributor, or a transposed table fails these.

def _interp_np(kd, okd, atm, col_dry):
    """Independent numpy restatement of the interpolation state (interpolation_kernel,
    gas_optics_rrtmgp_kernels.cu:245-277) per (col, lay): jtemp, ftemp, itropo, locpress and
    eta_0/eta_1 per flavor."""
    tref = kd.temp_ref
    ntemp, dT = tref.size, tref[1] - tref[0]
    T = atm.t_lay
    jtemp = np.clip(((T - (tref[0] - dT)) / dT).astype(int), 1, ntemp - 1)
    ftemp = (T - tref[jtemp - 1]) / dT
    lnp = np.log(atm.p_lay)
    lpr = np.log(kd.press_ref)
    dlp = (lpr[-1] - lpr[0]) / (lpr.size - 1)
    locp = 1.0 + (lnp - lpr[0]) / dlp
    itropo = (~(lnp > np.log(kd.press_ref_trop))).astype(int)
    names = okd.gas_names
    col = [col_dry] + [np.broadcast_to(atm.vmr[n], col_dry.shape) * col_dry for n in names]
    vref = okd.get_real("vmr_ref").reshape(ntemp, len(names) + 1, 2)     # oracle copy, (t, gas, tropo)
    flav = okd.get_int("flavor").reshape(-1, 2)
    eta = np.zeros((flav.shape[0], 2) + T.shape)
    for f, (g1, g2) in enumerate(flav):
        for i in (0, 1):
            r = vref[jtemp - 1 + i, g1, itropo] / vref[jtemp - 1 + i, g2, itropo]
            cm = col[g1] + r * col[g2]
            eta[f, i] = np.where(cm > 2 * np.finfo(float).tiny, col[g1] / cm, 0.5)
    return dict(jtemp=jtemp, ftemp=ftemp, itropo=itropo, locp=locp, eta=eta, T=T)


def _eta_nodes(kd):
    neta = kd.kmajor.shape[1]
    return np.arange(neta) / (neta - 1.0)


def test_interp_np_matches_oracle_indices(O, kd_lw, okd_lw, atm_c0, col_dry):
    # The numpy restatement agrees with the oracle's materialised indices, so the closed forms
    # below rest on the same cells.
    s = _interp_np(kd_lw, okd_lw, atm_c0, col_dry)
    o = O.gas_optics(okd_lw, atm_c0, col_dry=col_dry, interp=True)
    assert (o["jtemp"] == s["jtemp"]).all()
    assert (o["tropo"].astype(int) == 1 - s["itropo"]).all()
    neta = kd_lw.kmajor.shape[1]
    jeta = np.minimum((s["eta"] * (neta - 1)).astype(int) + 1, neta - 1)
    # oracle jeta is (itemp, col, lay, flavor); the restatement's is (flavor, itemp, col, lay)
    assert (o["jeta"] == jeta.transpose(1, 2, 3, 0)).all()


def test_planck_fraction_closed_form(O, kd_lw, okd_lw, atm_c0, col_dry):
    ngpt, neta, npp, ntemp = kd_lw.planck_frac.shape
    c = 1.0 + 0.01 * np.arange(ngpt)
    a, b, d = 0.7, 0.9, 0.3
    tab = c[:, None, None, None] * (1.0 + a * kd_lw.temp_ref[None, None, None, :] / 300.0
                                    + b * _eta_nodes(kd_lw)[None, :, None, None]
                                    + d * (np.arange(npp) / npp)[None, None, :, None])
    kd = dataclasses.replace(kd_lw, planck_frac=np.asfortranarray(tab),
                             totplnk=np.asfortranarray(np.ones_like(kd_lw.totplnk)))
    okd = O.Kdist(kd, AVAILABLE)
    o = O.gas_optics(okd, atm_c0, col_dry=col_dry)
    s = _interp_np(kd, okd, atm_c0, col_dry)
    gf = okd.get_int("gpoint_flavor").reshape(-1, 2)
    expect = np.zeros_like(o["lay_src"])
    for g in range(ngpt):
        f = gf[g, s["itropo"]] - 1                       # flavor of g in each cell's half
        e0 = np.take_along_axis(s["eta"][:, 0], f[None], 0)[0]
        e1 = np.take_along_axis(s["eta"][:, 1], f[None], 0)[0]
        eta_t = (1 - s["ftemp"]) * e0 + s["ftemp"] * e1
        # pressure node (0-based) of the lower corner is jpress - 1 + itropo
        pn = s["locp"] - 1 + s["itropo"]
        expect[:, :, g] = c[g] * (1 + a * s["T"] / 300.0 + b * eta_t + d * pn / npp)
    assert rel_err(o["lay_src"], expect) < 1e-13
    assert rel_err(o["lev_src_inc"], expect) < 1e-13 and rel_err(o["lev_src_dec"], expect) < 1e-13
    # surface source: the fraction of the surface layer (layer 1 when p decreases upward)
    sfc = 0 if atm_c0.p_lay[0, 0] > atm_c0.p_lay[0, -1] else -1
    assert rel_err(o["sfc_src"], expect[:, sfc, :]) < 1e-13


def _minor_scaling(kd, atm, col_dry, side, i):
    vmr = {k: np.broadcast_to(v, col_dry.shape) for k, v in atm.vmr.items()}
    gas_of = dict(zip(kd.identifier_minor, kd.gas_minor))
    s = vmr[gas_of[getattr(kd, f"minor_gases_{side}")[i]]] * col_dry
    if getattr(kd, f"minor_scales_with_density_{side}")[i]:
        s = s * atm.p_lay / 100.0 / atm.t_lay
        sg = getattr(kd, f"scaling_gas_{side}")[i]
        if sg:
            frac = vmr[sg] / (1.0 + vmr["h2o"])
            s = s * ((1.0 - frac) if getattr(kd, f"scale_by_complement_{side}")[i] else frac)
    return s


def test_kminor_closed_form(O, kd_lw, okd_lw, atm_c0, col_dry):
    slopes = {"lower": (0.8, 1.3), "upper": (-0.4, 2.1)}
    ch = {"kmajor": np.asfortranarray(np.zeros_like(kd_lw.kmajor))}
    for side, (a, b) in slopes.items():
        km = getattr(kd_lw, f"kminor_{side}")                # (ncontrib, neta, ntemp)
        cc = 1.0 + 0.02 * np.arange(km.shape[0])
        ch[f"kminor_{side}"] = np.asfortranarray(
            cc[:, None, None] * (1.0 + a * kd_lw.temp_ref[None, None, :] / 300.0
                                 + b * _eta_nodes(kd_lw)[None, :, None]))
    kd = dataclasses.replace(kd_lw, **ch)
    okd = O.Kdist(kd, AVAILABLE)
    tau = O.gas_optics(okd, atm_c0, col_dry=col_dry)["tau"]
    s = _interp_np(kd, okd, atm_c0, col_dry)
    gf = okd.get_int("gpoint_flavor").reshape(-1, 2)
    expect = np.zeros_like(tau)
    for side, itr in (("lower", 0), ("upper", 1)):
        a, b = slopes[side]
        mask = s["itropo"] == itr
        lim = getattr(kd, f"minor_limits_gpt_{side}")
        start = getattr(kd, f"kminor_start_{side}")
        cc = 1.0 + 0.02 * np.arange(getattr(kd, f"kminor_{side}").shape[0])
        for i in range(lim.shape[1]):
            g0, g1 = lim[:, i]
            f = gf[g0 - 1, itr] - 1                          # flavor of the interval's first g-point
            eta_t = (1 - s["ftemp"]) * s["eta"][f, 0] + s["ftemp"] * s["eta"][f, 1]
            base = (1 + a * s["T"] / 300.0 + b * eta_t) * _minor_scaling(kd, atm_c0, col_dry, side, i)
            for g in range(g0 - 1, g1):
                k = start[i] - 1 + (g - (g0 - 1))             # contributor of g-point g
                expect[:, :, g] += np.where(mask, cc[k] * base, 0.0)
    assert rel_err(tau, expect, floor=1e-300) < 1e-13


def test_rayleigh_closed_form_lower_upper(O, kd_sw, okd_sw, atm_c0, col_dry):
    ngpt = kd_sw.ngpt
    par = {"lower": (0.6, 1.7, 1e-27), "upper": (-0.3, 0.5, 3e-27)}
    ch = {"kmajor": np.asfortranarray(np.zeros_like(kd_sw.kmajor)),
          "kminor_lower": np.asfortranarray(np.zeros_like(kd_sw.kminor_lower)),
          "kminor_upper": np.asfortranarray(np.zeros_like(kd_sw.kminor_upper))}
    cg = 1.0 + 0.03 * np.arange(ngpt)
    for side, (a, b, k0) in par.items():
        ch[f"rayl_{side}"] = np.asfortranarray(
            k0 * cg[:, None, None] * (1.0 + a * kd_sw.temp_ref[None, None, :] / 300.0
                                      + b * _eta_nodes(kd_sw)[None, :, None]))
    kd = dataclasses.replace(kd_sw, **ch)
    okd = O.Kdist(kd, AVAILABLE)
    o = O.gas_optics(okd, atm_c0, col_dry=col_dry)
    s = _interp_np(kd, okd, atm_c0, col_dry)
    gf = okd.get_int("gpoint_flavor").reshape(-1, 2)
    col = atm_c0.vmr["h2o"] * col_dry + col_dry
    expect = np.zeros_like(o["tau"])
    for g in range(ngpt):
        for side, itr in (("lower", 0), ("upper", 1)):
            a, b, k0 = par[side]
            f = gf[g, itr] - 1
            eta_t = (1 - s["ftemp"]) * s["eta"][f, 0] + s["ftemp"] * s["eta"][f, 1]
            val = k0 * cg[g] * (1 + a * s["T"] / 300.0 + b * eta_t) * col
            expect[:, :, g] = np.where(s["itropo"] == itr, val, expect[:, :, g])
    assert rel_err(o["tau"], expect) < 1e-13
    assert np.all(o["ssa"] == 1.0)
