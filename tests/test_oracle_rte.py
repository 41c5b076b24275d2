"""Pins the oracle's cloud optics, optical-property arithmetic, LW/SW solvers, flux reductions
and heating rates with closed forms, limits and an independent dense linear solve."""
import numpy as np
import pytest

from conftest import rel_err

PI = np.pi


def _band_lims(nbnd, gpb):
    return np.array([np.arange(nbnd) * gpb + 1, (np.arange(nbnd) + 1) * gpb], dtype=np.int32)


# ----------------------------------------------------------------------------------------- LW
def _lw_case(rng, ncol=5, nlay=7, ngpt=6, nbnd=2):
    tau = rng.uniform(0.0, 2.0, (ncol, nlay, ngpt))
    lay = rng.uniform(1.0, 5.0, (ncol, nlay, ngpt))
    linc = rng.uniform(1.0, 5.0, (ncol, nlay, ngpt))
    ldec = rng.uniform(1.0, 5.0, (ncol, nlay, ngpt))
    sfc = rng.uniform(1.0, 5.0, (ncol, ngpt))
    emis = rng.uniform(0.8, 1.0, (nbnd, ncol))
    return tau, lay, linc, ldec, sfc, emis, _band_lims(nbnd, ngpt // nbnd)


def test_lw_transparent(O):
    rng = np.random.default_rng(1)
    tau, lay, linc, ldec, sfc, emis, bl = _lw_case(rng)
    up, dn = O.rte_lw(0 * tau, lay, linc, ldec, sfc, emis, bl, top_at_1=True)
    e = np.repeat(emis.T, 3, axis=1)
    assert np.all(dn == 0.0)
    assert rel_err(up, np.broadcast_to((PI * e * sfc)[:, None, :], up.shape)) < 1e-14


@pytest.mark.parametrize("nmu", [1, 2, 3, 4])
def test_lw_isothermal(O, nmu):
    # B everywhere, emissivity 1: up = pi*B at every level; dn = 2*pi*sum_i w_i B (1 - exp(-D_i tau_above)).
    rng = np.random.default_rng(2)
    tau, *_ , bl = _lw_case(rng)
    B = 3.0
    ones = np.ones_like(tau)
    ncol, nlay, ngpt = tau.shape
    up, dn = O.rte_lw(tau, B * ones, B * ones, B * ones, B * np.ones((ncol, ngpt)),
                      np.ones((2, ncol)), bl, top_at_1=True, n_gauss_angles=nmu)
    Ds = {1: [1.66], 2: [1.18350343, 2.81649655], 3: [1.09719858, 1.69338507, 4.70941630],
          4: [1.06056257, 1.38282560, 2.40148179, 7.15513024]}[nmu]
    ws = {1: [0.5], 2: [0.3180413817, 0.1819586183], 3: [0.2009319137, 0.2292411064, 0.0698269799],
          4: [0.1355069134, 0.2034645680, 0.1298475476, 0.0311809710]}[nmu]
    tau_above = np.concatenate([np.zeros((ncol, 1, ngpt)), np.cumsum(tau, axis=1)], axis=1)
    expect_dn = sum(2 * PI * w * B * (1 - np.exp(-D * tau_above)) for D, w in zip(Ds, ws))
    assert rel_err(dn, expect_dn, floor=1e-3) < 1e-12
    assert rel_err(up, 2 * PI * sum(ws) * B * np.ones_like(up)) < 1e-12


@pytest.mark.parametrize("tau0", [1e-10, 1e-5, 0.1, 1.0, 10.0])
def test_lw_single_layer_linear_in_tau(O, tau0):
    # With lay = (B_top + B_bot)/2 the linear-in-tau source equals the exact integral of
    # dI/dt = B(t) - I for B linear in optical depth (Clough et al. 1992, Eq. 13).
    Bt, Bb = 2.0, 5.0
    D = 1.66
    t = D * tau0
    T = np.exp(-t)
    a = -np.expm1(-t) / t - T          # (1-T)/t - T, accurate for small t
    exact_dn = Bb * (1 - T - a) + Bt * a
    exact_up_from_sfc = Bt * (1 - T - a) + Bb * a
    arr = lambda v: np.full((1, 1, 1), v)
    sfc = 4.0
    up, dn = O.rte_lw(arr(tau0), arr(0.5 * (Bt + Bb)), arr(Bb), arr(Bt), np.full((1, 1), sfc),
                      np.ones((1, 1)), np.array([[1], [1]], dtype=np.int32), top_at_1=True)
    # Above the sqrt(eps) switch the reference evaluates (1-T)/t - T directly, which carries an
    # absolute error ~eps/t; that conditioning term is part of the tolerance.
    tol = 1e-12 * exact_dn + 8 * np.finfo(float).eps / t * abs(Bb - Bt) + 1e-15
    assert abs(dn[0, 1, 0] / PI - exact_dn) <= tol
    expect_up_top = sfc * T + exact_up_from_sfc        # emissivity 1: up_sfc = sfc source
    assert abs(up[0, 0, 0] / PI - expect_up_top) <= 1e-12 * expect_up_top + tol


def test_lw_orientation_and_broadband(O):
    rng = np.random.default_rng(3)
    tau, lay, linc, ldec, sfc, emis, bl = _lw_case(rng)
    up, dn = O.rte_lw(tau, lay, linc, ldec, sfc, emis, bl, top_at_1=True)
    # Flipped column: inc/dec sources swap roles.
    f = lambda a: a[:, ::-1]
    up2, dn2 = O.rte_lw(f(tau), f(lay), f(ldec), f(linc), sfc, emis, bl, top_at_1=False)
    assert rel_err(up2[:, ::-1], up) < 1e-14 and rel_err(dn2[:, ::-1], dn, floor=1e-12) < 1e-14
    bu, bd = O.rte_lw(tau, lay, linc, ldec, sfc, emis, bl, top_at_1=True, do_broadband=True)
    assert rel_err(bu, up.sum(2)) < 1e-13 and rel_err(bd, dn.sum(2), floor=1e-12) < 1e-13


def test_lw_jacobian_is_transmitted_surface_source(O):
    rng = np.random.default_rng(4)
    tau, lay, linc, ldec, sfc, emis, bl = _lw_case(rng)
    jac_src = rng.uniform(0.01, 0.1, sfc.shape)
    _, _, jac = O.rte_lw(tau, lay, linc, ldec, sfc, emis, bl, top_at_1=True, sfc_src_jac=jac_src,
                         do_broadband=True)
    z = np.zeros_like(tau)
    up, _ = O.rte_lw(tau, z, z, z, jac_src, emis, bl, top_at_1=True, do_broadband=True)
    assert rel_err(jac, up) < 1e-13


# ----------------------------------------------------------------------------------------- SW
def _sw_run(O, tau, ssa, g, mu0, S, alb_dir, alb_dif, top_at_1=True, bl=None):
    ncol, nlay, ngpt = tau.shape
    if bl is None:
        bl = np.array([[1], [ngpt]], dtype=np.int32)
    inc = np.full((ncol, ngpt), S)
    return O.rte_sw(tau, ssa, g, mu0, inc, np.full((bl.shape[1], ncol), alb_dir),
                    np.full((bl.shape[1], ncol), alb_dif), bl, top_at_1)


def test_sw_transparent(O):
    ncol, nlay = 3, 9
    mu0 = np.array([0.2, 0.5, 1.0])
    z = np.zeros((ncol, nlay, 1))
    up, dn, dr = _sw_run(O, z, z, z, mu0, 1000.0, 0.3, 0.3)
    m = (mu0 * 1000.0)[:, None, None]
    assert rel_err(dr, np.broadcast_to(m, dr.shape)) < 1e-14
    assert rel_err(dn, np.broadcast_to(m, dn.shape)) < 1e-12
    assert rel_err(up, np.broadcast_to(0.3 * m, up.shape)) < 1e-12


def test_sw_no_scattering_beer_lambert(O):
    # omega = 0: PIFM gives R_dif = 0, T_dif = exp(-2 tau); the beam follows Beer-Lambert and the
    # surface-reflected diffuse light is attenuated with diffusivity 2 on the way up.
    rng = np.random.default_rng(5)
    ncol, nlay = 4, 8
    tau = rng.uniform(0, 0.5, (ncol, nlay, 1))
    mu0 = rng.uniform(0.1, 1.0, ncol)
    z = np.zeros_like(tau)
    up, dn, dr = _sw_run(O, tau, z, z, mu0, 1000.0, 0.25, 0.25)
    above = np.concatenate([np.zeros((ncol, 1, 1)), np.cumsum(tau, 1)], 1)
    below = above[:, -1:, :] - above
    dir_exact = (mu0 * 1000.0)[:, None, None] * np.exp(-above / mu0[:, None, None])
    up_exact = 0.25 * dir_exact[:, -1:, :] * np.exp(-2 * below)
    assert rel_err(dr, dir_exact) < 1e-13
    # The Rdir/Tdir clamp to [eps, ...] adds eps*dir per layer: an absolute floor of ~nlay*eps*S.
    atol = 16 * nlay * np.finfo(float).eps * 1000.0
    assert np.max(np.abs(dn - dir_exact) - 1e-12 * dir_exact) < atol
    assert np.max(np.abs(up - up_exact) - 1e-12 * up_exact) < atol


def test_sw_conservative_scattering_conserves_energy(O):
    rng = np.random.default_rng(6)
    ncol, nlay = 4, 10
    tau = rng.uniform(0, 3.0, (ncol, nlay, 1))
    mu0 = rng.uniform(0.2, 1.0, ncol)
    up, dn, dr = _sw_run(O, tau, np.ones_like(tau), 0.85 * np.ones_like(tau), mu0, 1000.0, 0.2, 0.2)
    net = dn - up
    # No absorption in the atmosphere: net flux constant with height; surface absorbs (1-albedo).
    assert np.max(np.abs(net - net[:, :1])) < 1e-4 * 1000.0
    assert rel_err(net[:, -1], 0.8 * dn[:, -1]) < 1e-12      # top_at_1: surface is the last level


def test_sw_two_stream_limits(O):
    tau = np.array([0.3, 1.0, 50.0, 2.0, 1e-3])
    w0 = np.array([0.0, 0.0, 0.9, 1.0, 0.5])
    g = np.array([0.0, 0.5, 0.5, 0.0, 0.7])
    mu0 = np.array([0.5, 0.8, 0.6, 0.7, 0.3])
    Rdif, Tdif, Rdir, Tdir, Tns = O.sw_two_stream(tau, w0, g, mu0)
    eps = np.finfo(float).eps
    np.testing.assert_allclose(Tns, np.exp(-tau / mu0), rtol=1e-15)
    # No scattering: R_dif = 0, T_dif = exp(-2 tau) (gamma1 = 2, k = 2), direct sources clamp to eps.
    assert Rdif[0] == 0.0 and Rdif[1] == 0.0
    np.testing.assert_allclose(Tdif[:2], np.exp(-2 * tau[:2]), rtol=1e-15)
    assert Rdir[0] == eps and Tdir[0] == eps
    # Semi-infinite layer: R_dif -> gamma2 / (gamma1 + k) (PIFM, Zdunkowski et al. 1980).
    g1 = (8 - w0[2] * (5 + 3 * g[2])) / 4
    g2 = 3 * w0[2] * (1 - g[2]) / 4
    k = np.sqrt(g1 ** 2 - g2 ** 2)
    assert abs(Rdif[2] - g2 / (g1 + k)) < 1e-15
    assert rel_err(Tdif[2], 2 * k * np.exp(-k * tau[2]) / (k + g1)) < 1e-13
    # Conservative scattering: nothing is absorbed (up to the k_min = 1e-12 floor).
    assert abs(Rdif[3] + Tdif[3] - 1) < 1e-5
    assert abs(Rdir[3] + Tdir[3] + Tns[3] - 1) < 1e-5
    # Everything stays physical: R, T in [0, 1] and R_dir + T_dir <= 1 - T_noscat.
    assert np.all(Rdir + Tdir <= 1 - Tns + 1e-15) and np.all(Rdif >= 0) and np.all(Tdif <= 1)


@pytest.mark.parametrize("top_at_1", [True, False])
def test_sw_adding_matches_dense_solve(O, top_at_1):
    """The adding method must reproduce the solution of the two-stream layer equations."""
    rng = np.random.default_rng(7)
    nlay = 6
    tau = rng.uniform(0.01, 2.0, nlay)
    w0 = rng.uniform(0.2, 0.99, nlay)
    g = rng.uniform(0.0, 0.9, nlay)
    mu0, S, a_dir, a_dif = 0.6, 1000.0, 0.3, 0.15
    Rd, Td, Rb, Tb, Tn = O.sw_two_stream(tau, w0, g, np.full(nlay, mu0))
    # Levels ordered top (0) to surface (nlay); unknowns x = [up_0..up_n, dn_0..dn_n] (diffuse).
    n = nlay + 1
    dirf = np.zeros(n)
    dirf[0] = mu0 * S
    for l in range(nlay):
        dirf[l + 1] = Tn[l] * dirf[l]
    A = np.zeros((2 * n, 2 * n))
    b = np.zeros(2 * n)
    UP, DN = (lambda k: k), (lambda k: n + k)
    r = 0
    A[r, DN(0)] = 1.0; r += 1                                   # no diffuse flux at TOA
    for l in range(nlay):
        A[r, UP(l)] = 1; A[r, DN(l)] = -Rd[l]; A[r, UP(l + 1)] = -Td[l]; b[r] = Rb[l] * dirf[l]; r += 1
        A[r, DN(l + 1)] = 1; A[r, DN(l)] = -Td[l]; A[r, UP(l + 1)] = -Rd[l]; b[r] = Tb[l] * dirf[l]; r += 1
    A[r, UP(nlay)] = 1; A[r, DN(nlay)] = -a_dif; b[r] = a_dir * dirf[nlay]
    x = np.linalg.solve(A, b)
    up_x, dn_x = x[:n], x[n:] + dirf
    arr = lambda v: v[None, :, None]
    if top_at_1:
        up, dn, dr = _sw_run(O, arr(tau), arr(w0), arr(g), np.array([mu0]), S, a_dir, a_dif, True)
        up, dn, dr = up[0, :, 0], dn[0, :, 0], dr[0, :, 0]
    else:
        up, dn, dr = _sw_run(O, arr(tau[::-1]), arr(w0[::-1]), arr(g[::-1]), np.array([mu0]), S,
                             a_dir, a_dif, False)
        up, dn, dr = up[0, ::-1, 0], dn[0, ::-1, 0], dr[0, ::-1, 0]
    assert rel_err(dr, dirf) < 1e-14
    assert rel_err(up, up_x) < 1e-12 and rel_err(dn, dn_x) < 1e-12


def test_sw_broadband_and_orientation(O):
    rng = np.random.default_rng(8)
    ncol, nlay, ngpt = 3, 6, 4
    tau = rng.uniform(0, 1, (ncol, nlay, ngpt)); ssa = rng.uniform(0, 1, tau.shape); g = rng.uniform(0, 0.9, tau.shape)
    mu0 = rng.uniform(0.1, 1, ncol); inc = rng.uniform(10, 100, (ncol, ngpt))
    bl = _band_lims(2, 2)
    ad, af = rng.uniform(0, 0.5, (2, ncol)), rng.uniform(0, 0.5, (2, ncol))
    up, dn, dr = O.rte_sw(tau, ssa, g, mu0, inc, ad, af, bl, True)
    bu, bd, bdr = O.rte_sw(tau, ssa, g, mu0, inc, ad, af, bl, True, do_broadband=True)
    assert rel_err(bu, up.sum(2)) < 1e-13 and rel_err(bd, dn.sum(2)) < 1e-13 and rel_err(bdr, dr.sum(2)) < 1e-13
    f = lambda a: a[:, ::-1]
    up2, dn2, dr2 = O.rte_sw(f(tau), f(ssa), f(g), mu0, inc, ad, af, bl, False)
    assert rel_err(f(up2), up) < 1e-13 and rel_err(f(dn2), dn) < 1e-13


# ------------------------------------------------------------------------------ cloud optics
def test_cloud_nodes_and_combination(O, lut_sw):
    cl = O.Cloud(lut_sw)
    nb = lut_sw.nbnd
    step = (lut_sw.radliq_upr - lut_sw.radliq_lwr) / 19
    istep = (lut_sw.radice_upr - lut_sw.radice_lwr) / 19
    nodes = [0, 5, 19]
    rel = np.array([[lut_sw.radliq_lwr + k * step for k in nodes] + [0.0]])
    lwp = np.array([[10.0, 20.0, 30.0, 0.0]])
    rei = np.array([[0.0, 0.0, lut_sw.radice_lwr + 7 * istep, lut_sw.radice_lwr + 7.5 * istep]])
    iwp = np.array([[0.0, 0.0, 5.0, 8.0]])
    tau, ssa, g = cl.optics(lwp, iwp, rel, rei, True)
    E, W, G = lut_sw.lut_extliq, lut_sw.lut_ssaliq, lut_sw.lut_asyliq
    Ei, Wi, Gi = (a[:, :, 1] for a in (lut_sw.lut_extice, lut_sw.lut_ssaice, lut_sw.lut_asyice))
    for j, k in enumerate(nodes[:2]):
        assert rel_err(tau[0, j], lwp[0, j] * E[k]) < 1e-13
        assert rel_err(ssa[0, j], W[k]) < 1e-13 and rel_err(g[0, j], G[k]) < 1e-13
    # liquid at the last node + ice at node 7 (middle roughness table)
    tl, ti = 30.0 * E[19], 5.0 * Ei[7]
    assert rel_err(tau[0, 2], tl + ti) < 1e-13
    assert rel_err(ssa[0, 2], (tl * W[19] + ti * Wi[7]) / (tl + ti)) < 1e-13
    assert rel_err(g[0, 2], (tl * W[19] * G[19] + ti * Wi[7] * Gi[7]) / (tl * W[19] + ti * Wi[7])) < 1e-13
    # ice half-way between nodes 7 and 8: linear interpolation of each table
    ext = 0.5 * (Ei[7] + Ei[8]); om = 0.5 * (Wi[7] + Wi[8])
    assert rel_err(tau[0, 3], 8.0 * ext) < 1e-13 and rel_err(ssa[0, 3], om) < 1e-13
    tau1 = cl.optics(lwp, iwp, rel, rei, False)
    assert rel_err(tau1[0, 0], 10.0 * E[0] * (1 - W[0])) < 1e-13
    assert rel_err(tau1[0, 2], tl * (1 - W[19]) + ti * (1 - Wi[7])) < 1e-13
    assert nb == tau.shape[2]


def test_cloud_zero_water_is_clear(O, lut_lw):
    cl = O.Cloud(lut_lw)
    z = np.zeros((3, 4))
    assert np.all(cl.optics(z, z, z, z, False) == 0.0)


# -------------------------------------------------------------------- optical-props arithmetic
def test_delta_scale_invariants(O):
    rng = np.random.default_rng(9)
    tau, ssa, g = rng.uniform(0, 5, (3, 4, 5)), rng.uniform(0, 1, (3, 4, 5)), rng.uniform(0, 0.95, (3, 4, 5))
    t2, s2, g2 = O.delta_scale(tau, ssa, g)
    assert rel_err(t2 * (1 - s2), tau * (1 - ssa), floor=1e-12) < 1e-12      # absorption kept
    f = g * g
    assert rel_err(g2, g / (1 + g)) < 1e-13                                  # (g - g^2)/(1 - g^2)
    assert rel_err(t2 * s2, tau * ssa * (1 - f), floor=1e-12) < 1e-12        # scattering removed
    z = np.zeros_like(g)
    t3, s3, g3 = O.delta_scale(tau, ssa, z)
    assert np.all(t3 == tau) and np.all(s3 == ssa) and np.all(g3 == 0)


def test_add_to_bybnd_conserves_components(O):
    rng = np.random.default_rng(10)
    ncol, nlay, nbnd, gpb = 3, 4, 3, 2
    ngpt = nbnd * gpb
    bl = _band_lims(nbnd, gpb)
    t1, s1, g1 = rng.uniform(0, 1, (ncol, nlay, ngpt)), rng.uniform(0, 1, (ncol, nlay, ngpt)), rng.uniform(0, 0.9, (ncol, nlay, ngpt))
    t2, s2, g2 = rng.uniform(0, 1, (ncol, nlay, nbnd)), rng.uniform(0, 1, (ncol, nlay, nbnd)), rng.uniform(0, 0.9, (ncol, nlay, nbnd))
    t, s, g = O.inc_bybnd(t1, t2, bl, s1, g1, s2, g2)
    rep = lambda a: np.repeat(a, gpb, axis=2)
    assert rel_err(t, t1 + rep(t2)) < 1e-14
    assert rel_err(t * s, t1 * s1 + rep(t2 * s2)) < 1e-13
    assert rel_err(t * s * g, t1 * s1 * g1 + rep(t2 * s2 * g2)) < 1e-13
    t_only = O.inc_bybnd(t1, t2, bl)
    assert rel_err(t_only, t1 + rep(t2)) < 1e-14
    ts, ss, gs = O.increment(t1, t1, s1, g1, s1, g1)
    assert rel_err(ts, 2 * t1) < 1e-15 and rel_err(ss, s1) < 1e-14 and rel_err(gs, g1) < 1e-13


# ----------------------------------------------------------------------------- reductions, HR
def test_flux_reductions(O):
    rng = np.random.default_rng(11)
    f = rng.uniform(0, 10, (4, 6, 12))
    bl = np.array([[1, 4, 9], [3, 8, 12]], dtype=np.int32)
    assert rel_err(O.sum_broadband(f), f.sum(2)) < 1e-14
    bb = O.sum_byband(f, bl)
    assert rel_err(bb[:, :, 0], f[:, :, 0:3].sum(2)) < 1e-14
    assert rel_err(bb[:, :, 1], f[:, :, 3:8].sum(2)) < 1e-14
    assert rel_err(bb[:, :, 2], f[:, :, 8:12].sum(2)) < 1e-14


def test_heating_rate_linear_profile(O):
    # net flux linear in pressure -> uniform heating rate -(g/cp) dF/dp, in K/day
    p = np.linspace(1e5, 1e3, 11)[None, :].repeat(2, 0)
    net = 100.0 - 2e-3 * p
    hr = O.heating_rate(net, p)
    expect = 9.80665 / 1004.64 * 2e-3 * 86400.0
    assert rel_err(hr, np.full_like(hr, expect)) < 1e-10
    hr2 = O.heating_rate(net[:, ::-1], p[:, ::-1])
    assert rel_err(hr2, hr[:, ::-1]) < 1e-12


# -------------------------------------------------------------------------------- drivers
def test_solve_lw_block_size_independent_and_composed(O, okd_lw, syn, lut_lw):
    atm = syn.make_atmosphere(21, 17, 40, cloudy=True)
    cl = O.Cloud(lut_lw)
    a = O.solve_lw(okd_lw, atm, cl, n_col_block=12)
    b = O.solve_lw(okd_lw, atm, cl, n_col_block=5)
    for k in a:
        assert np.array_equal(a[k], b[k])
    c = O.solve_lw(okd_lw, atm, cl, bnd=True)
    assert rel_err(c["flux_up"], a["flux_up"]) < 1e-12
    assert rel_err(c["bnd_flux_dn"].sum(2), c["flux_dn"], floor=1e-9) < 1e-12
    # composition of the class-level steps
    go = O.gas_optics(okd_lw, atm)
    tau = O.inc_bybnd(go["tau"], cl.optics(atm.lwp, atm.iwp, atm.rel, atm.rei, False), okd_lw.band_lims)
    up, dn = O.rte_lw(tau, go["lay_src"], go["lev_src_inc"], go["lev_src_dec"], go["sfc_src"],
                      atm.emis_sfc, okd_lw.band_lims, atm.top_at_1)
    assert rel_err(up.sum(2), a["flux_up"]) < 1e-12 and rel_err(dn.sum(2), a["flux_dn"], floor=1e-9) < 1e-12


def test_solve_sw_composed(O, okd_sw, syn, lut_sw):
    atm = syn.make_atmosphere(22, 13, 40, cloudy=True)
    cl = O.Cloud(lut_sw)
    a = O.solve_sw(okd_sw, atm, cl)
    go = O.gas_optics(okd_sw, atm)
    ct, cs, cg = O.delta_scale(*cl.optics(atm.lwp, atm.iwp, atm.rel, atm.rei, True))
    tau, ssa, g = O.inc_bybnd(go["tau"], ct, okd_sw.band_lims, go["ssa"], go["g"], cs, cg)
    toa = go["toa_src"] * (atm.tsi / okd_sw.scalar("tsi"))
    up, dn, dr = O.rte_sw(tau, ssa, g, atm.mu0, toa, atm.sfc_alb, atm.sfc_alb, okd_sw.band_lims, atm.top_at_1)
    assert rel_err(up.sum(2), a["flux_up"]) < 1e-12
    assert rel_err(dn.sum(2), a["flux_dn"]) < 1e-12
    assert rel_err(dr.sum(2), a["flux_dn_dir"], floor=1e-9) < 1e-12
    np.testing.assert_allclose(a["flux_dn"][:, -1], atm.mu0 * atm.tsi, rtol=1e-12)


def test_solve_threads_identical(O, okd_lw, okd_sw, lut_lw, lut_sw, syn):
    """The OpenMP block loop gives the serial loop's results bit for bit, errors included."""
    atm = syn.make_atmosphere(41, 30, 40, cloudy=True)
    tsi = np.full(atm.ncol, atm.tsi / okd_sw.scalar("tsi"))
    prev = O.set_threads(1)
    try:
        a = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw)), O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), tsi_scaling=tsi)
        O.set_threads(4)
        b = O.solve_lw(okd_lw, atm, O.Cloud(lut_lw)), O.solve_sw(okd_sw, atm, O.Cloud(lut_sw), tsi_scaling=tsi)
        for x, y in zip(a, b):
            for n in x:
                assert np.array_equal(x[n], y[n]), n
        bad = syn.make_atmosphere(41, 30, 40)
        bad.t_lay[17, 5] = 400.0
        with pytest.raises(ValueError):
            O.solve_lw(okd_lw, bad)
    finally:
        O.set_threads(prev)


# --------------------------------------------------------------------- LW rescaling, LW 2-stream
def test_lw_rescaling_reduces_to_noscat(O):
    """Without scattering the rescaled solver is the no-scattering solver; with g = 1 (Cn = 0) it is
    the no-scattering solver on the scaled optical depth tau (1 - ssa)."""
    rng = np.random.default_rng(11)
    tau, lay, linc, ldec, sfc, emis, bl = _lw_case(rng)
    z = np.zeros_like(tau)
    for nmu in (1, 3):
        up, dn = O.rte_lw(tau, lay, linc, ldec, sfc, emis, bl, top_at_1=True, n_gauss_angles=nmu)
        up2, dn2 = O.rte_lw_rescaled(tau, z, rng.uniform(0, 1, tau.shape), lay, linc, ldec, sfc, emis, bl,
                                     top_at_1=True, n_gauss_angles=nmu)
        assert np.array_equal(up, up2) and np.array_equal(dn, dn2)
        ssa = rng.uniform(0, 1, tau.shape)
        for top_at_1 in (True, False):
            up, dn = O.rte_lw(tau * (1 - ssa), lay, linc, ldec, sfc, emis, bl, top_at_1=top_at_1, n_gauss_angles=nmu)
            up2, dn2 = O.rte_lw_rescaled(tau, ssa, np.ones_like(tau), lay, linc, ldec, sfc, emis, bl,
                                         top_at_1=top_at_1, n_gauss_angles=nmu)
            assert rel_err(up2, up) < 1e-13 and rel_err(dn2, dn, floor=1e-12) < 1e-13


@pytest.mark.parametrize("top_at_1", [True, False])
def test_lw_rescaling_single_layer_reflection(O, top_at_1):
    """One non-emitting layer over a black surface lit from above.  With g = -1 the optical depth is
    not scaled and Cn = 0.4 ssa: the up pass reflects Cn An of the incident radiance at the layer
    top (An = 1 - T^2); the second down pass adds Cn An times the up radiance of level ilev, which
    is the layer's top (the reflected radiance) when top_at_1 and its bottom (zero) otherwise."""
    tau0, w, F = 0.7, 0.6, 50.0
    T = np.exp(-1.66 * tau0)
    An, Cn = 1 - T * T, 0.4 * w
    arr = lambda v: np.full((1, 1, 1), v)
    z = arr(0.0)
    up, dn = O.rte_lw_rescaled(arr(tau0), arr(w), arr(-1.0), z, z, z, np.zeros((1, 1)), np.ones((1, 1)),
                               np.array([[1], [1]], dtype=np.int32), top_at_1, inc_flux=np.full((1, 1), F))
    top, bot = (0, 1) if top_at_1 else (1, 0)
    assert abs(up[0, top, 0] - Cn * An * F) < 1e-13 * F and up[0, bot, 0] == 0.0
    want_dn = T * F + (Cn * An * Cn * An * F if top_at_1 else 0.0)
    assert abs(dn[0, bot, 0] - want_dn) < 1e-13 * F and dn[0, top, 0] == F


def test_lw_rescaling_jacobian_and_broadband(O):
    """The surface-temperature Jacobian is transported without the adjustment, i.e. as the
    no-scattering Jacobian on the scaled optical depth; broadband fluxes are the g-point sums."""
    rng = np.random.default_rng(12)
    tau, lay, linc, ldec, sfc, emis, bl = _lw_case(rng)
    ssa, g = rng.uniform(0, 1, tau.shape), rng.uniform(0, 0.9, tau.shape)
    jac_src = rng.uniform(0.01, 0.1, sfc.shape)
    scale = 1 - ssa + ssa * (1 - g) * 0.5
    bu, bd, jac = O.rte_lw_rescaled(tau, ssa, g, lay, linc, ldec, sfc, emis, bl, True, do_broadband=True,
                                    sfc_src_jac=jac_src)
    _, _, jac0 = O.rte_lw(tau * scale, lay, linc, ldec, sfc, emis, bl, True, do_broadband=True, sfc_src_jac=jac_src)
    assert rel_err(jac, jac0) < 1e-13
    up, dn = O.rte_lw_rescaled(tau, ssa, g, lay, linc, ldec, sfc, emis, bl, True)
    assert rel_err(bu, up.sum(2)) < 1e-13 and rel_err(bd, dn.sum(2), floor=1e-12) < 1e-13


def _lw2(O, tau, ssa, g, lev_top, lev_bot, emis, sfc, top_at_1=True, inc=None):
    """lw_solver_2stream with level sources given per layer top/bottom (a column-constant source
    profile combines to those values exactly: sqrt(B B) = B)."""
    ncol, nlay, ngpt = tau.shape
    return O.lw_solver_2stream(tau, ssa, g, lev_bot if top_at_1 else lev_top, lev_top if top_at_1 else lev_bot,
                               emis, sfc, top_at_1, inc_flux=inc)


def test_lw_2stream_isothermal_equilibrium(O):
    """An isothermal column at B over a black surface at B, lit from above by pi B: up = dn = pi B on
    every level, whatever the scattering (thermodynamic equilibrium); without scattering and without
    incident flux, dn = pi B (1 - exp(-1.66 tau_above))."""
    rng = np.random.default_rng(13)
    ncol, nlay, ngpt = 3, 6, 4
    tau = rng.uniform(0.0, 2.0, (ncol, nlay, ngpt))
    ssa, g = rng.uniform(0, 1, tau.shape), rng.uniform(0, 0.9, tau.shape)
    B = 3.0
    src = np.full(tau.shape, B)
    one, sfc = np.ones((ncol, ngpt)), np.full((ncol, ngpt), B)
    for top_at_1 in (True, False):
        up, dn = O.lw_solver_2stream(tau, ssa, g, src, src, one, sfc, top_at_1, inc_flux=np.full((ncol, ngpt), PI * B))
        assert rel_err(up, np.full(up.shape, PI * B)) < 1e-13 and rel_err(dn, np.full(dn.shape, PI * B)) < 1e-13
    up, dn = O.lw_solver_2stream(tau, 0 * ssa, g, src, src, one, sfc, True)
    tau_above = np.concatenate([np.zeros((ncol, 1, ngpt)), np.cumsum(tau, axis=1)], axis=1)
    assert rel_err(up, np.full(up.shape, PI * B)) < 1e-13
    assert rel_err(dn, PI * B * (1 - np.exp(-1.66 * tau_above)), floor=1e-3) < 1e-12


@pytest.mark.parametrize("tau0", [1e-4, 0.1, 1.0, 8.0])
@pytest.mark.parametrize("top_at_1", [True, False])
def test_lw_2stream_single_layer_linear_source(O, tau0, top_at_1):
    """No scattering, one layer with a Planck source linear in optical depth between the level
    values: the emitted fluxes equal the quadrature of the diffusivity-approximation integrals
    pi int_0^tau B(t) 1.66 exp(-1.66 t) dt (up, at the top) and with exp(-1.66 (tau - t)) (down,
    at the bottom)."""
    from scipy.integrate import quad
    Bt, Bb, Bs = 2.0, 5.0, 4.0
    B = lambda t: Bt + (Bb - Bt) * t / tau0
    e_up = quad(lambda t: B(t) * 1.66 * np.exp(-1.66 * t), 0, tau0, epsabs=0, epsrel=1e-13)[0]
    e_dn = quad(lambda t: B(t) * 1.66 * np.exp(-1.66 * (tau0 - t)), 0, tau0, epsabs=0, epsrel=1e-13)[0]
    arr = lambda v: np.full((1, 1, 1), v)
    up, dn = _lw2(O, arr(tau0), arr(0.0), arr(0.5), arr(Bt), arr(Bb), np.ones((1, 1)), np.full((1, 1), Bs), top_at_1)
    top, bot = (0, 1) if top_at_1 else (1, 0)
    T = np.exp(-1.66 * tau0)
    assert abs(up[0, top, 0] - PI * (e_up + T * Bs)) < 1e-11 * PI * Bb
    assert abs(dn[0, bot, 0] - PI * e_dn) < 1e-11 * PI * Bb
    assert up[0, bot, 0] == PI * Bs and dn[0, top, 0] == 0.0


def test_lw_2stream_layer_doubling(O):
    """Two-stream reflectance and transmittance of homogeneous layers obey the doubling rule: a
    layer of 2 tau reflects and transmits as two layers of tau combined (Rdif, Tdif of the
    Meador-Weaver solution); with no emission and a black surface, up(top) = R F, dn(sfc) = T F."""
    rng = np.random.default_rng(14)
    n = 6
    tau, w, gg = rng.uniform(0.05, 3.0, n), rng.uniform(0.0, 1.0, n), rng.uniform(0.0, 0.9, n)
    F = 100.0
    z2 = np.zeros((n, 2, 1))
    one, zero, inc = np.ones((n, 1)), np.zeros((n, 1)), np.full((n, 1), F)
    col = lambda a, k: np.repeat(a[:, None, None], k, axis=1)
    up1, dn1 = O.lw_solver_2stream(col(2 * tau, 1), col(w, 1), col(gg, 1), z2[:, :1], z2[:, :1], one, zero, True, inc)
    up2, dn2 = O.lw_solver_2stream(col(tau, 2), col(w, 2), col(gg, 2), z2, z2, one, zero, True, inc)
    assert rel_err(up1[:, 0, 0], up2[:, 0, 0]) < 1e-12 and rel_err(dn1[:, 1, 0], dn2[:, 2, 0]) < 1e-12
    # and the semi-infinite limit R = gamma2 / (gamma1 + k) of the Fu et al. coefficients
    g1 = 1.66 * (1 - 0.5 * w * (1 + gg))
    g2 = 1.66 * 0.5 * w * (1 - gg)
    k = np.sqrt(np.maximum((g1 - g2) * (g1 + g2), 1e-12))
    upi, _ = O.lw_solver_2stream(col(np.full(n, 60.0), 1), col(w, 1), col(gg, 1), z2[:, :1], z2[:, :1], one, zero,
                                 True, inc)
    assert rel_err(upi[:, 0, 0], F * g2 / (g1 + k), floor=1e-30) < 1e-12


def test_lw_2stream_conservative_scattering_conserves_energy(O):
    """ssa = 1 and no emission: the net flux is constant with height (to the k_min floor)."""
    rng = np.random.default_rng(15)
    ncol, nlay, ngpt = 2, 8, 3
    tau = rng.uniform(0.1, 3.0, (ncol, nlay, ngpt))
    g = rng.uniform(0, 0.9, tau.shape)
    z = np.zeros_like(tau)
    emis = np.full((ncol, ngpt), 0.6)
    up, dn = O.lw_solver_2stream(tau, np.ones_like(tau), g, z, z, emis, np.zeros((ncol, ngpt)), True,
                                 inc_flux=np.full((ncol, ngpt), 200.0))
    net = dn - up
    assert np.all(net[:, -1] > 10.0)
    assert rel_err(net, np.broadcast_to(net[:, -1:, :], net.shape)) < 1e-5


@pytest.mark.parametrize("top_at_1", [True, False])
def test_lw_2stream_level_sources_combine(O, top_at_1):
    """No scattering, three layers whose edge sources come from one level profile (lev_inc of a
    layer = the profile at its increasing-index level, lev_dec at its decreasing-index level): each
    interior level combines to the profile value, so the top up and surface down fluxes are the
    layer emissions, each the quadrature of its linear-in-tau integral, transmitted through the
    layers between."""
    from scipy.integrate import quad
    Bl = np.array([1.0, 2.5, 6.0, 3.0])      # level profile by increasing index
    tau = np.array([0.3, 1.2, 0.5])
    Bs = 4.0
    nlay = 3
    # Order the layers top to bottom for the integrals.
    order = range(nlay) if top_at_1 else range(nlay - 1, -1, -1)
    lay_tb = [(Bl[l], Bl[l + 1]) if top_at_1 else (Bl[l + 1], Bl[l]) for l in order]
    t_tb = [tau[l] for l in order]

    def emit(bt, bb, t0, up):
        B = lambda t: bt + (bb - bt) * t / t0
        w = (lambda t: np.exp(-1.66 * t)) if up else (lambda t: np.exp(-1.66 * (t0 - t)))
        return quad(lambda t: B(t) * 1.66 * w(t), 0, t0, epsabs=0, epsrel=1e-13)[0]

    T = np.exp(-1.66 * np.array(t_tb))
    up_top = Bs * np.prod(T) + sum(emit(*lay_tb[k], t_tb[k], True) * np.prod(T[:k]) for k in range(nlay))
    dn_sfc = sum(emit(*lay_tb[k], t_tb[k], False) * np.prod(T[k + 1:]) for k in range(nlay))
    arr = lambda v: np.asarray(v, float)[None, :, None]
    up, dn = O.lw_solver_2stream(arr(tau), arr(np.zeros(3)), arr(np.full(3, 0.5)), arr(Bl[1:]), arr(Bl[:-1]),
                                 np.ones((1, 1)), np.full((1, 1), Bs), top_at_1)
    top, sfc = (0, nlay) if top_at_1 else (nlay, 0)
    assert abs(up[0, top, 0] - PI * up_top) < 1e-11 * PI * 6.0
    assert abs(dn[0, sfc, 0] - PI * dn_sfc) < 1e-11 * PI * 6.0
