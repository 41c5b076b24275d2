"""Pins of the oracle's aerosol optics (src/Aerosol_optics.cpp:26-224) on the reference's own
table (data/aerosol_optics.nc, kept as tests/golden/aerosol_optics.nc).

The species -> (mixing-ratio field, table, column) assignment below is the one the reference's
compute_all_from_table spells out (:62-139); each species is lit alone so that a swapped field,
a wrong column, a hydrophobic/hydrophilic mix-up or a transposed (band, humidity, species) index
changes tau, ssa or g by the table's spread.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err

GOLD = os.path.join(ROOT, "tests", "golden")

# field number -> (hydrophilic?, 1-based table column), as compute_all_from_table assigns them.
SPECIES = {1: (True, 1), 2: (True, 2), 3: (True, 3), 4: (False, 1), 5: (False, 8), 6: (False, 6),
           7: (True, 4), 8: (False, 10), 9: (False, 11), 10: (False, 11), 11: (True, 5)}


@pytest.fixture(scope="module")
def lut():
    from rrtmgp_amd import coefficients, netcdf
    return coefficients.aerosol_lut(netcdf.Dataset(os.path.join(GOLD, "aerosol_optics.nc")))


def _zero_fields(ncol, nlay):
    return {f"aermr{i:02d}": np.zeros((ncol, nlay)) for i in range(1, 12)}


def _levels(ncol, nlay, rng):
    plev = np.sort(rng.uniform(100.0, 1.0e5, (ncol, nlay + 1)), axis=1)[:, ::-1]     # surface first
    return np.asfortranarray(plev)


def test_single_species_closed_form(O, lut):
    """One species alone: tau = mmr |dp| / 9.81 * mext, ssa and g the table's, per band."""
    rng = np.random.default_rng(5)
    ncol, nlay = 6, 4
    plev = _levels(ncol, nlay, rng)
    dpg = (plev[:, :-1] - plev[:, 1:]) / 9.81
    nhum = lut["rh_upper"].size
    ihum = rng.integers(0, nhum, (ncol, nlay))
    rh = lut["rh_upper"][ihum] - 1e-3            # inside class ihum: the bounds are >= 0.05 apart
    for field, (philic, column) in SPECIES.items():
        mmr = rng.uniform(1e-9, 1e-7, (ncol, nlay))
        f = _zero_fields(ncol, nlay)
        f[f"aermr{field:02d}"] = mmr
        tau, ssa, g = O.aerosol_optics(f, rh, plev, lut)
        for b in range(tau.shape[2]):
            if philic:
                mext = lut["mext_philic"][b, ihum, column - 1]
                s, gg = lut["ssa_philic"][b, ihum, column - 1], lut["g_philic"][b, ihum, column - 1]
            else:
                mext = np.full((ncol, nlay), lut["mext_phobic"][b, column - 1])
                s = np.full((ncol, nlay), lut["ssa_phobic"][b, column - 1])
                gg = np.full((ncol, nlay), lut["g_phobic"][b, column - 1])
            assert rel_err(tau[:, :, b], mmr * dpg * mext) < 1e-14, (field, b)
            assert rel_err(ssa[:, :, b], s) < 1e-14, (field, b)
            assert rel_err(g[:, :, b], gg) < 1e-13, (field, b)


def test_humidity_class_bounds(O, lut):
    """A relative humidity on a class's upper bound belongs to that class, just above it to the
    next; below the first bound is class 1, above the last the last class."""
    ru = lut["rh_upper"]
    rh = np.array([[ru[0] * 0.5, ru[0], ru[0] * (1 + 1e-9), ru[4], ru[4] + 1e-9, ru[-1], ru[-1] + 0.3]])
    want = np.array([0, 0, 1, 4, 5, len(ru) - 1, len(ru) - 1])
    ncol, nlay = rh.shape
    plev = np.array([[1.0e5, 9.0e4, 8.0e4, 7.0e4, 6.0e4, 5.0e4, 4.0e4, 3.0e4]])
    f = _zero_fields(ncol, nlay)
    f["aermr01"] = np.full((ncol, nlay), 1e-8)                  # SS1: hydrophilic column 1
    tau, ssa, g = O.aerosol_optics(f, rh, plev, lut)
    for b in range(tau.shape[2]):
        assert rel_err(ssa[0, :, b], lut["ssa_philic"][b, want, 0]) < 1e-15, b


def test_mixture_and_zero(O, lut):
    """Two species: tau adds, ssa and g are the scattering-weighted means; no aerosol gives zeros."""
    rng = np.random.default_rng(6)
    ncol, nlay = 5, 3
    plev = _levels(ncol, nlay, rng)
    rh = rng.uniform(0, 1, (ncol, nlay))
    f = _zero_fields(ncol, nlay)
    tau0, ssa0, g0 = O.aerosol_optics(f, rh, plev, lut)
    assert not tau0.any() and not ssa0.any() and not g0.any()
    f1, f2 = _zero_fields(ncol, nlay), _zero_fields(ncol, nlay)
    f1["aermr05"] = rng.uniform(1e-9, 1e-7, (ncol, nlay))      # dust, hydrophobic
    f2["aermr11"] = rng.uniform(1e-9, 1e-7, (ncol, nlay))      # sulphate, hydrophilic
    t1, s1, g1 = O.aerosol_optics(f1, rh, plev, lut)
    t2, s2, g2 = O.aerosol_optics(f2, rh, plev, lut)
    f12 = dict(f1)
    f12["aermr11"] = f2["aermr11"]
    t, s, g = O.aerosol_optics(f12, rh, plev, lut)
    assert rel_err(t, t1 + t2) < 1e-14
    assert rel_err(s, (t1 * s1 + t2 * s2) / (t1 + t2)) < 1e-13
    assert rel_err(g, (t1 * s1 * g1 + t2 * s2 * g2) / (t1 * s1 + t2 * s2)) < 1e-13


def test_profile_broadcast_and_orientation(O, lut):
    """(1, nlay) profiles equal their (ncol, nlay) expansion; the layer mass is |dp|, so flipping the
    level order (with the layers) flips the result."""
    rng = np.random.default_rng(7)
    ncol, nlay = 4, 6
    plev = _levels(ncol, nlay, rng)
    rh = rng.uniform(0, 1, (ncol, nlay))
    prof = {f"aermr{i:02d}": rng.uniform(0, 1e-7, (1, nlay)) for i in range(1, 12)}
    full = {k: np.repeat(v, ncol, axis=0) for k, v in prof.items()}
    a = O.aerosol_optics(prof, rh, plev, lut)
    b = O.aerosol_optics(full, rh, plev, lut)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    flip = O.aerosol_optics({k: v[:, ::-1] for k, v in full.items()}, rh[:, ::-1], plev[:, ::-1], lut)
    for x, y in zip(flip, b):
        assert np.array_equal(x, y[:, ::-1, :])
