"""Write the NetCDF-4 (HDF5) fixtures of tests/test_netcdf.py with an independent HDF5 library.

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9):

    /opt/conda/bin/python3.9 tests/golden/make_netcdf_fixtures.py

Each file lays netCDF-4's conventions over HDF5 (dimension-scale datasets, DIMENSION_LIST
references, an unnamed dimension written as a scale without data, creation-order tracking) and
exercises one storage form the reader must handle:

  earliest.nc   superblock 0, symbol-table root group, contiguous / compact / chunked storage
                (version-1 B-tree chunk index), deflate + shuffle + Fletcher-32, big-endian,
                fill values for unwritten chunks, scalar and char variables
  tracked.nc    as netCDF-C writes: creation order tracked, so links and attributes beyond eight
                are stored densely (fractal heap + version-2 B-tree)
  many.nc       300 variables in one tracked group: a multi-level version-2 B-tree and a
                fractal heap with indirect blocks; var_<i>_... holds [i, i / 2, -i]
  latest.nc     superblock 3, layout version 4 (single-chunk, fixed-array and implicit indices)

Expected values go to expected_<name>.npz (numpy arrays, no pickles) and the dimension names
and attributes to expected.json.  A reading of data/aerosol_optics.nc (from the reference's
tree, when present) by the same library is stored as per-variable checksums in
expected_aerosol.json.
"""
import json
import os
import sys

import h5py
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "netcdf")
NOT_A_VAR = "This is a netCDF dimension but not a netCDF variable.%10d"


def dims_of(f, spec):
    for name, size, has_data in spec:
        if has_data:
            d = f.create_dataset(name, data=np.arange(size, dtype="f8") * 0.5 + 1.0)
            d.make_scale(name)
        else:   # netCDF-C names the scale of a dimension without a variable this way
            d = f.create_dataset(name, shape=(size,), dtype="f4")
            d.make_scale(NOT_A_VAR % size)


def attach(f, var, names):
    for i, n in enumerate(names):
        f[var].dims[i].attach_scale(f[n])


def common(f, rng, dense_attrs, kw_chunk):
    out, meta = {}, {}
    dims_of(f, [("x", 5, True), ("y", 3, False), ("lay", 4, True), ("string_len", 8, False),
                ("pair", 2, False)])
    a = rng.standard_normal((4, 3, 5))
    f.create_dataset("p_lay", data=a)
    attach(f, "p_lay", ("lay", "y", "x"))
    out["p_lay"], meta["p_lay"] = a, ["lay", "y", "x"]

    b = rng.standard_normal((4, 3, 5)).astype("f4")
    f.create_dataset("t_lay", data=b, chunks=(2, 2, 3), compression="gzip", compression_opts=4,
                     shuffle=True, **kw_chunk)
    attach(f, "t_lay", ("lay", "y", "x"))
    out["t_lay"], meta["t_lay"] = b, ["lay", "y", "x"]

    c = rng.integers(-1000, 1000, size=(3, 5)).astype(">i4")
    f.create_dataset("key_species", data=c, chunks=(2, 5), fletcher32=True, **kw_chunk)
    attach(f, "key_species", ("y", "x"))
    out["key_species"], meta["key_species"] = c.astype("<i4"), ["y", "x"]

    f.create_dataset("press_ref_trop", data=np.float64(9948.431564193395))
    out["press_ref_trop"], meta["press_ref_trop"] = np.array(9948.431564193395), []

    names = ["h2o", "co2", "o3"]
    s = np.full((3, 8), b" ", dtype="S1")
    for i, n in enumerate(names):
        s[i, :len(n)] = np.frombuffer(n.encode(), dtype="S1")
    dims_of(f, [("absorber", 3, False)])
    f.create_dataset("gas_names", data=s)
    attach(f, "gas_names", ("absorber", "string_len"))
    out["gas_names"], meta["gas_names"] = s, ["absorber", "string_len"]

    e = np.array([[1, -2], [3, -4], [5, 6]], dtype="i1")
    f.create_dataset("bnd_flags", data=e)
    attach(f, "bnd_flags", ("absorber", "pair"))
    out["bnd_flags"], meta["bnd_flags"] = e, ["absorber", "pair"]

    g = np.linspace(0, 1, 10).reshape(2, 5)
    sid = h5py.h5s.create_simple((2, 5))
    dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
    dcpl.set_layout(h5py.h5d.COMPACT)
    h5py.h5d.create(f.id, b"compact_var", h5py.h5t.IEEE_F64LE, sid, dcpl).write(
        h5py.h5s.ALL, h5py.h5s.ALL, g)
    attach(f, "compact_var", ("pair", "x"))
    out["compact_var"], meta["compact_var"] = g, ["pair", "x"]

    # Partly written chunked variable: unwritten chunks read as the fill value.
    h = f.create_dataset("sparse", shape=(4, 5), chunks=(2, 2), dtype="f8", fillvalue=-7.5,
                         compression="gzip", **kw_chunk)
    h[0:2, 0:2] = 3.25
    attach(f, "sparse", ("lay", "x"))
    hv = np.full((4, 5), -7.5)
    hv[0:2, 0:2] = 3.25
    out["sparse"], meta["sparse"] = hv, ["lay", "x"]

    v = f["t_lay"]
    v.attrs["units"] = np.bytes_("K")
    v.attrs["scale"] = np.float32(2.5)
    v.attrs["valid_range"] = np.array([100., 400.])
    for i in range(dense_attrs):
        v.attrs[f"extra_{i:02d}"] = np.int32(i * 3)
    w = f.create_dataset("many_attrs", data=np.arange(3.0), track_order=True)
    attach(f, "many_attrs", ("absorber",))
    for i in range(20):
        w.attrs[f"a{i:02d}"] = np.float64(i) / 4
    out["many_attrs"], meta["many_attrs"] = np.arange(3.0), ["absorber"]
    return out, meta


def main():
    os.makedirs(HERE, exist_ok=True)
    expected = {}
    rng = np.random.default_rng(20261203)

    cases = [
        ("earliest", dict(libver="earliest"), 2, {}),
        ("tracked", dict(libver="earliest", track_order=True), 12, {}),
        ("latest", dict(libver="latest", track_order=True), 12, {}),
    ]
    for name, kw, nattr, kw_chunk in cases:
        path = os.path.join(HERE, f"{name}.nc")
        with h5py.File(path, "w", **kw) as f:
            out, meta = common(f, rng, nattr, kw_chunk)
            if name == "latest":
                one = rng.standard_normal((3, 5))
                f.create_dataset("single_chunk", data=one, chunks=(3, 5), compression="gzip")
                attach(f, "single_chunk", ("y", "x"))
                out["single_chunk"], meta["single_chunk"] = one, ["y", "x"]
                imp = rng.standard_normal((4, 5))
                sid = h5py.h5s.create_simple((4, 5))
                dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
                dcpl.set_chunk((2, 5))
                dcpl.set_alloc_time(h5py.h5d.ALLOC_TIME_EARLY)
                h5py.h5d.create(f.id, b"implicit_idx", h5py.h5t.IEEE_F64LE, sid, dcpl).write(
                    h5py.h5s.ALL, h5py.h5s.ALL, imp)
                attach(f, "implicit_idx", ("lay", "x"))
                out["implicit_idx"], meta["implicit_idx"] = imp, ["lay", "x"]
        np.savez(os.path.join(HERE, f"expected_{name}.npz"), **out)
        expected[name] = {"dims": meta,
                          "attrs_t_lay": {"units": "K", "scale": 2.5, "valid_range": [100., 400.],
                                          **{f"extra_{i:02d}": i * 3 for i in range(nattr)}},
                          "dimensions": {"x": 5, "y": 3, "lay": 4, "string_len": 8, "pair": 2,
                                         "absorber": 3}}

    path = os.path.join(HERE, "many.nc")
    many = {}
    with h5py.File(path, "w", libver="earliest", track_order=True) as f:
        dims_of(f, [("n", 3, False)])
        for i in range(300):
            nm = f"var_{i:03d}_{'abcdefghij' * (i % 4)}"
            val = np.array([i, i * 0.5, -i], dtype="f8")
            f.create_dataset(nm, data=val)
            attach(f, nm, ("n",))
            many[nm] = val
    expected["many"] = {"count": len(many), "values": "var_<i>_* holds [i, i / 2, -i]"}

    with open(os.path.join(HERE, "expected.json"), "w") as fh:
        json.dump(expected, fh, indent=1, sort_keys=True)

    ref = "/root/reference/data/aerosol_optics.nc"
    if os.path.exists(ref):
        summary = {}
        with h5py.File(ref, "r") as f:
            for nm, d in f.items():
                if not isinstance(d, h5py.Dataset):
                    continue
                a = d[()]
                entry = {"shape": list(a.shape), "dtype": a.dtype.str}
                dl = d.attrs.get("DIMENSION_LIST")
                if dl is not None:
                    entry["dims"] = [f[r[0]].name.lstrip("/") for r in dl]
                if a.dtype.kind in "fiu":
                    x = a.astype("f8")
                    entry["sum"] = float(np.nansum(x))
                    entry["sumsq"] = float(np.nansum(x * x))
                    entry["first"] = float(x.reshape(-1)[0]) if x.size else None
                    entry["last"] = float(x.reshape(-1)[-1]) if x.size else None
                elif a.dtype.kind == "S":
                    entry["text"] = b"".join(a.reshape(-1).tolist()).decode()
                summary[nm] = entry
        with open(os.path.join(HERE, "expected_aerosol.json"), "w") as fh:
            json.dump(summary, fh, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
