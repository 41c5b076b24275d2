"""The NetCDF reader (rrtmgp_amd.netcdf) against files written and read by an independent HDF5
library (tests/golden/make_netcdf_fixtures.py, h5py), the reference's data/aerosol_optics.nc
(copied to tests/golden/aerosol_optics.nc) and NetCDF-3 files written by scipy.
"""
import json
import os

import numpy as np
import pytest

from rrtmgp_amd import coefficients, netcdf

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NC = os.path.join(GOLD, "netcdf")


def _expected():
    with open(os.path.join(NC, "expected.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["earliest", "tracked", "latest"])
def test_values_dims_and_attributes(name):
    exp = _expected()[name]
    ds = netcdf.Dataset(os.path.join(NC, f"{name}.nc"))
    assert ds.dimensions == exp["dimensions"]
    ref = np.load(os.path.join(NC, f"expected_{name}.npz"))
    for k in ref.files:
        v = ds.variables[k]
        assert list(v.dimensions) == exp["dims"][k], k
        a = v.data
        assert a.shape == ref[k].shape, k
        assert a.dtype.kind == ref[k].dtype.kind and a.dtype.itemsize == ref[k].dtype.itemsize, k
        assert np.array_equal(a, ref[k]), k
    # Dimensions without a variable are dimensions only; coordinate variables are both.
    assert "y" not in ds.variables and "pair" not in ds.variables
    assert np.array_equal(ds.variables["x"].data, np.arange(5) * 0.5 + 1.0)
    attrs = ds.variables["t_lay"].attributes
    for k, v in exp["attrs_t_lay"].items():
        assert np.allclose(attrs[k], v) if not isinstance(v, str) else attrs[k] == v, k
    many = ds.variables["many_attrs"].attributes
    assert sorted(many) == [f"a{i:02d}" for i in range(20)]
    assert all(float(many[f"a{i:02d}"]) == i / 4 for i in range(20))


def test_unwritten_chunks_read_as_fill_value():
    ds = netcdf.Dataset(os.path.join(NC, "tracked.nc"))
    a = ds.get_variable("sparse", (4, 5))
    assert (a[:2, :2] == 3.25).all() and (a[2:, :] == -7.5).all() and (a[:, 2:] == -7.5).all()


def test_many_links_multilevel_btree():
    ds = netcdf.Dataset(os.path.join(NC, "many.nc"))
    assert ds.dimensions == {"n": 3}
    assert len(ds.variables) == _expected()["many"]["count"]
    for k, v in ds.variables.items():
        i = int(k[4:7])
        assert v.dimensions == ("n",)
        assert np.array_equal(v.data, [i, i / 2, -i]), k


def test_aerosol_file_matches_independent_reader():
    """Every variable of the reference's aerosol table, against h5py's reading of the same file."""
    with open(os.path.join(NC, "expected_aerosol.json")) as f:
        exp = json.load(f)
    ds = netcdf.Dataset(os.path.join(GOLD, "aerosol_optics.nc"))
    assert ds.dimensions == {"wavelength_mono": 20, "code": 2, "hydrophobic": 14, "band_lw": 16,
                             "relative_humidity": 12, "hydrophilic": 7, "band_sw": 14}
    for name, e in exp.items():
        if name in ds.dimensions and name not in ds.variables:
            assert ds.dimensions[name] == e["shape"][0]
            continue
        v = ds.variables[name]
        a = v.data
        assert list(a.shape) == e["shape"], name
        if "dims" in e:
            assert list(v.dimensions) == e["dims"], name
        if "text" in e:
            assert b"".join(a.reshape(-1).tolist()).decode() == e["text"]
            continue
        x = a.astype("f8")
        assert float(np.nansum(x)) == pytest.approx(e["sum"], rel=1e-12, abs=1e-300), name
        assert float(np.nansum(x * x)) == pytest.approx(e["sumsq"], rel=1e-12, abs=1e-300), name
        assert float(x.reshape(-1)[0]) == e["first"] and float(x.reshape(-1)[-1]) == e["last"], name


def test_aerosol_reference_loader_shapes():
    """The reads of load_and_init_aerosol_optics (src_test/Radiation_solver.cpp:329-366)."""
    ds = netcdf.Dataset(os.path.join(GOLD, "aerosol_optics.nc"))
    n_band = ds.get_dimension_size("band_sw")
    n_hum = ds.get_dimension_size("relative_humidity")
    n_philic = ds.get_dimension_size("hydrophilic")
    n_phobic = ds.get_dimension_size("hydrophobic")
    mext = ds.get_variable("mass_ext_sw_hydrophilic", (n_philic, n_hum, n_band))
    ssa = ds.get_variable("ssa_sw_hydrophobic", (n_phobic, n_band))
    rh2 = ds.get_variable("relative_humidity2", (n_hum,))
    rh1 = ds.get_variable("relative_humidity1", (n_hum,))
    assert mext.shape == (7, 12, 14) and ssa.shape == (14, 14)
    # Bin edges chain and bands tile the spectrum: a transposed or shifted read breaks these.
    assert np.array_equal(rh1[1:], rh2[:-1])
    w1, w2 = ds.get_variable("wavenumber1_sw"), ds.get_variable("wavenumber2_sw")
    assert np.all(w2 > w1)
    assert ((ssa >= 0) & (ssa <= 1)).all()
    assert ds.get_strings("code_hydrophilic")[:3] == ["SS", "SS", "SS"]
    with pytest.raises(ValueError):
        ds.get_variable("mass_ext_sw_hydrophilic", (7, 12, 15))
    with pytest.raises(KeyError):
        ds.get_variable("no_such_variable")


def test_bundle_round_trip(tmp_path):
    ds = netcdf.Dataset(os.path.join(GOLD, "aerosol_optics.nc"))
    out = str(tmp_path / "aerosol_bundle")
    assert netcdf.main([os.path.join(GOLD, "aerosol_optics.nc"), out]) == 0
    s = coefficients.read_bundle(out)
    for name, v in ds.variables.items():
        assert s.vars[name][0] == tuple(v.dimensions)
        assert np.array_equal(s.get(name), v.data), name


def test_classic_netcdf3(tmp_path):
    import scipy.io
    path = str(tmp_path / "classic.nc")
    f = scipy.io.netcdf_file(path, "w", version=2)
    f.createDimension("lay", 3)
    f.createDimension("x", 4)
    v = f.createVariable("t_lay", "f8", ("lay", "x"))
    v[:] = np.arange(12.0).reshape(3, 4)
    v.units = "K"
    s = f.createVariable("tsi_scaling", "f4", ())
    s.data[()] = 0.5
    f.close()
    ds = netcdf.Dataset(path)
    assert ds.dimensions == {"lay": 3, "x": 4}
    assert ds.variables["t_lay"].dimensions == ("lay", "x")
    assert np.array_equal(ds.get_variable("t_lay", (3, 4)), np.arange(12.0).reshape(3, 4))
    assert float(ds.get_variable("tsi_scaling")) == 0.5
    assert ds.variables["t_lay"].attributes["units"] == "K"


def test_not_netcdf(tmp_path):
    p = tmp_path / "junk.nc"
    p.write_bytes(b"this is not a netcdf file at all" * 10)
    with pytest.raises(netcdf.FormatError):
        netcdf.Dataset(str(p))


@pytest.mark.parametrize("version", [2, 5])
def test_writer_round_trip(tmp_path, version):
    """Writer output read back by this module's NetCDF-3 reader; CDF-2 also by scipy."""
    path = str(tmp_path / f"out{version}.nc")
    w = netcdf.Writer(path, version=version)
    w.add_dimension("x", 5)
    w.add_dimension("y", 1)
    w.add_dimension("lev", 4)
    w.add_dimension("pair", 2)
    up = np.random.default_rng(3).normal(size=(4, 1, 5))
    w.add_variable("lw_flux_up", ("lev", "y", "x"), up)
    w.add_variable("sw_flux_dn", ("lev", "y", "x"), up.astype(np.float32), {"units": "W m-2"})
    w.add_variable("lw_band_lims_gpt", ("x", "pair"), np.arange(10, dtype=np.int32).reshape(5, 2))
    w.add_variable("tsi", (), np.float64(1360.5))
    w.close()
    ds = netcdf.Dataset(path)
    assert ds.dimensions == {"x": 5, "y": 1, "lev": 4, "pair": 2}
    assert ds.variables["lw_flux_up"].dimensions == ("lev", "y", "x")
    assert np.array_equal(ds.get_variable("lw_flux_up"), up)
    assert ds.get_variable("sw_flux_dn").dtype == np.float32
    assert np.array_equal(ds.get_variable("sw_flux_dn"), up.astype(np.float32))
    assert ds.variables["sw_flux_dn"].attributes["units"] == "W m-2"
    assert np.array_equal(ds.get_variable("lw_band_lims_gpt"), np.arange(10).reshape(5, 2))
    assert float(ds.get_variable("tsi")) == 1360.5
    if version == 2:
        import scipy.io
        f = scipy.io.netcdf_file(path, "r", mmap=False)
        assert np.array_equal(f.variables["lw_flux_up"][:], up)
        assert f.variables["lw_band_lims_gpt"].dimensions == ("x", "pair")
        f.close()


def test_writer_checks(tmp_path):
    w = netcdf.Writer(str(tmp_path / "bad.nc"))
    w.add_dimension("x", 3)
    with pytest.raises(KeyError):
        w.add_variable("a", ("lay",), np.zeros(3))
    with pytest.raises(ValueError):
        w.add_variable("a", ("x",), np.zeros(4))
    with pytest.raises(ValueError):
        w.add_dimension("x", 4)
