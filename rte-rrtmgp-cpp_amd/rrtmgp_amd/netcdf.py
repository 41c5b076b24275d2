"""Read NetCDF files without netCDF-C: NetCDF-4 (HDF5) and NetCDF-3 classic.

The reference reads every coefficient set and case file through netCDF-C
(include_test/Netcdf_interface.h; Radiation_solver.cpp:70-366 for the k-distributions, cloud and
aerosol tables; test_rte_rrtmgp.cpp:205-293 for the case).  This image has neither netCDF-C nor
h5py, so this module reads the two on-disk formats those files come in:

* NetCDF-4, which is HDF5 with conventions (the RRTMGP data files and data/aerosol_optics.nc).
  Only what netCDF-4 writes is read: superblock versions 0-3, object headers version 1 and 2,
  groups stored as symbol tables or as link messages (compact), contiguous, compact and chunked
  storage (version-1 B-tree chunk index, and the single-chunk / fixed-array / implicit indices of
  layout version 4), the deflate, shuffle and Fletcher-32 filters, fixed-point, floating-point and
  fixed-length string types.  A variable's dimension names come from its DIMENSION_LIST attribute
  (object references to the dimension-scale datasets, held in the global heap).
* NetCDF-3: classic (CDF-1), 64-bit offset (CDF-2) and 64-bit data (CDF-5) files, read here
  without scipy.

`Writer` writes the NetCDF-3 formats (CDF-2, or CDF-5 once a fixed-size variable passes the
4 GiB that CDF-2 allows all but its last variable); the driver's output file
(test_rte_rrtmgp.cpp:332-611 writes it as NetCDF-4, with the same dimensions and variables) is
written through it.

`Dataset(path)` exposes `dimensions` (name -> size), `variables` (name -> `Variable`) and
`get_variable(name, dims)`, the counterpart of Netcdf_handle::get_variable with its size check.
`to_schema` turns a coefficient file into the `coefficients.Schema` the bundles are written from,
so a NetCDF coefficient file converts to the build's bundle format with

    python -m rrtmgp_amd.netcdf rrtmgp-gas-lw-g256.nc out_dir/
"""
from __future__ import annotations

import struct
import sys
import zlib

import numpy as np

_UNDEF = 0xFFFFFFFFFFFFFFFF


class FormatError(ValueError):
    """The file is not NetCDF, or uses an HDF5 feature that netCDF-4 does not write."""


# ---------------------------------------------------------------------------------------------
# HDF5 structures (HDF5 File Format Specification, versions 0-3 of the superblock).

class _Buf:
    def __init__(self, data: bytes, size_off: int = 8, size_len: int = 8):
        self.d = data
        self.so = size_off
        self.sl = size_len

    def u(self, pos: int, n: int) -> int:
        if pos < 0 or pos + n > len(self.d):
            raise FormatError(f"read of {n} bytes at {pos} runs past the end of the file")
        return int.from_bytes(self.d[pos:pos + n], "little")

    def off(self, pos: int) -> int:
        return self.u(pos, self.so)

    def length(self, pos: int) -> int:
        return self.u(pos, self.sl)

    def sig(self, pos: int, s: bytes):
        if self.d[pos:pos + len(s)] != s:
            raise FormatError(f"expected signature {s!r} at {pos}")

    def cstr(self, pos: int) -> str:
        end = self.d.index(b"\0", pos)
        return self.d[pos:end].decode("utf-8")


class _Datatype:
    """Datatype message: class, element size and the numpy dtype of the numeric classes."""

    def __init__(self, b: _Buf, pos: int):
        cv = b.u(pos, 1)
        self.cls = cv & 0x0F
        self.version = cv >> 4
        bits = b.u(pos + 1, 3)
        self.size = b.u(pos + 4, 4)
        self.dtype = None
        self.members = []       # compound: (name, offset, _Datatype)
        self.base = None        # variable-length / array: base type
        self.vlen_string = False
        order = ">" if bits & 1 else "<"
        if self.cls == 0:       # fixed point
            signed = bool(bits & 0x08)
            self.dtype = np.dtype(f"{order}{'i' if signed else 'u'}{self.size}")
        elif self.cls == 1:     # floating point
            if self.size not in (2, 4, 8):
                raise FormatError(f"floating-point size {self.size}")
            self.dtype = np.dtype(f"{order}f{self.size}")
        elif self.cls == 3:     # fixed-length string
            self.dtype = np.dtype(f"S{self.size}")
        elif self.cls == 6:     # compound
            nmemb = bits & 0xFFFF
            p = pos + 8
            for _ in range(nmemb):
                name = b.cstr(p)
                if self.version < 3:
                    p += (len(name) + 1 + 7) // 8 * 8
                else:
                    p += len(name) + 1
                if self.version == 1:
                    moff = b.u(p, 4)
                    p += 4 + 1 + 3 + 4 + 4 + 16   # offset, dimensionality, reserved, perm, reserved, dims
                elif self.version == 2:
                    moff = b.u(p, 4)
                    p += 4
                else:
                    nb = max(1, (self.size.bit_length() + 7) // 8)
                    moff = b.u(p, nb)
                    p += nb
                mt = _Datatype(b, p)
                p += mt.encoded_size
                self.members.append((name, moff, mt))
        elif self.cls == 7:     # reference
            pass
        elif self.cls == 9:     # variable length
            self.vlen_string = (bits & 0x3) == 1
            self.base = _Datatype(b, pos + 8)
        elif self.cls == 10:    # array
            p = pos + 8
            ndims = b.u(p, 1)
            p += 1 if self.version >= 3 else 4
            p += 4 * ndims
            if self.version < 3:
                p += 4 * ndims
            self.base = _Datatype(b, p)
        self.encoded_size = self._encoded_size(b, pos)

    def _encoded_size(self, b: _Buf, pos: int) -> int:
        # Size of the message itself, needed only where datatypes nest (compound members).
        if self.cls in (0, 1):
            return 8 + (4 if self.cls == 0 else 12)
        if self.cls in (3, 7):
            return 8
        if self.cls == 9:
            return 8 + self.base.encoded_size
        if self.cls == 6:
            p = pos + 8
            for name, _, mt in self.members:
                p += (len(name) + 1 + 7) // 8 * 8 if self.version < 3 else len(name) + 1
                if self.version == 1:
                    p += 32
                elif self.version == 2:
                    p += 4
                else:
                    p += max(1, (self.size.bit_length() + 7) // 8)
                p += mt.encoded_size
            return p - pos
        if self.cls == 10:
            ndims = b.u(pos + 8, 1)
            head = 8 + (1 if self.version >= 3 else 4) + 4 * ndims + (4 * ndims if self.version < 3 else 0)
            return head + self.base.encoded_size
        raise FormatError(f"datatype class {self.cls} is not used by netCDF-4 variables")


class _Object:
    """An object header's messages: (type, flags, data offset, data size)."""

    def __init__(self, f: "_Hdf5", addr: int):
        self.addr = addr
        self.messages = []
        b = f.b
        if b.d[addr:addr + 4] == b"OHDR":
            self._read_v2(b, addr)
        else:
            self._read_v1(b, addr)

    def _read_v1(self, b: _Buf, addr: int):
        if b.u(addr, 1) != 1:
            raise FormatError(f"object header version {b.u(addr, 1)} at {addr}")
        nmsg = b.u(addr + 2, 2)
        size = b.u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        while blocks and len(self.messages) < nmsg:
            start, size = blocks.pop(0)
            p, end = start, start + size
            while p + 8 <= end and len(self.messages) < nmsg:
                mtype, msize, mflags = b.u(p, 2), b.u(p + 2, 2), b.u(p + 4, 1)
                data = p + 8
                if mtype == 0x10:   # continuation
                    blocks.append((b.off(data), b.length(data + b.so)))
                self.messages.append((mtype, mflags, data, msize))
                p = data + msize

    def _read_v2(self, b: _Buf, addr: int):
        flags = b.u(addr + 5, 1)
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        n = 1 << (flags & 3)
        size0 = b.u(p, n)
        p += n
        blocks = [(p, size0, False)]
        tracked = bool(flags & 0x04)
        while blocks:
            start, size, cont = blocks.pop(0)
            q, end = start, start + size
            while q + 4 <= end:
                mtype, msize, mflags = b.u(q, 1), b.u(q + 1, 2), b.u(q + 3, 1)
                q += 4 + (2 if tracked else 0)
                if q + msize > end:
                    break
                if mtype == 0x10:
                    caddr, clen = b.off(q), b.length(q + b.so)
                    b.sig(caddr, b"OCHK")
                    blocks.append((caddr + 4, clen - 8, True))
                if mtype != 0:
                    self.messages.append((mtype, mflags, q, msize))
                q += msize

    def find(self, mtype: int):
        return [m for m in self.messages if m[0] == mtype]


def _nbytes(n: int) -> int:
    """Bytes needed to hold the integer n."""
    return (n.bit_length() + 7) // 8


class _FractalHeap:
    """Managed objects of a fractal heap (FRHP): where dense link and attribute storage lives.

    Managed space is the concatenation of the heap's direct blocks in the order of their heap
    offsets (each block counted from its own header), so an object's heap offset indexes it.
    """

    def __init__(self, f: "_Hdf5", addr: int):
        b = f.b
        b.sig(addr, b"FRHP")
        self.b = b
        self.id_len = b.u(addr + 5, 2)
        filt_len = b.u(addr + 7, 2)
        self.flags = b.u(addr + 9, 1)
        max_obj = b.u(addr + 10, 4)
        p = addr + 14 + b.sl + b.so + b.sl + b.so + 8 * b.sl
        self.width = b.u(p, 2)
        self.start = b.length(p + 2)
        self.max_direct = b.length(p + 2 + b.sl)
        max_heap_bits = b.u(p + 2 + 2 * b.sl, 2)
        root = b.off(p + 2 + 2 * b.sl + 4)
        root_rows = b.u(p + 2 + 2 * b.sl + 4 + b.so, 2)
        if filt_len:
            raise FormatError("filtered fractal heaps are not read")
        self.off_size = (max_heap_bits + 7) // 8
        self.len_size = _nbytes(min(self.max_direct, max_obj))
        self.heap_addr = addr
        self.max_direct_rows = (self.max_direct.bit_length() - 1) - (self.start.bit_length() - 1) + 2
        self.blocks = []          # (heap offset, file address, size)
        if root != _UNDEF:
            if root_rows == 0:
                self.blocks.append((0, root, self.start))
            else:
                self._indirect(root, root_rows)
        self.blocks.sort()

    def _row_size(self, r: int) -> int:
        return self.start if r == 0 else self.start << (r - 1)

    def _indirect(self, addr: int, nrows: int):
        b = self.b
        b.sig(addr, b"FHIB")
        block_off = b.u(addr + 5 + b.so, self.off_size)
        p = addr + 5 + b.so + self.off_size
        heap_off = block_off
        for r in range(nrows):
            size = self._row_size(r)
            for _ in range(self.width):
                child = b.off(p)
                p += b.so
                if r < self.max_direct_rows:
                    if child != _UNDEF:
                        self.blocks.append((heap_off, child, size))
                else:
                    if child != _UNDEF:
                        sub_rows = (size.bit_length() - 1) - ((self.start * self.width).bit_length() - 1) + 1
                        self._indirect(child, sub_rows)
                heap_off += size

    def get(self, heap_id: bytes) -> bytes:
        kind = (heap_id[0] >> 4) & 3
        if kind == 2:                     # tiny: the object is in the ID
            n = (heap_id[0] & 0x0F) + 1
            return bytes(heap_id[1:1 + n])
        if kind != 0:
            raise FormatError("huge fractal-heap objects are not read")
        off = int.from_bytes(heap_id[1:1 + self.off_size], "little")
        n = int.from_bytes(heap_id[1 + self.off_size:1 + self.off_size + self.len_size], "little")
        for hoff, addr, size in self.blocks:
            if hoff <= off < hoff + size:
                start = addr + (off - hoff)
                return bytes(self.b.d[start:start + n])
        raise FormatError(f"fractal-heap offset {off} lies in no direct block")


class _Hdf5:
    def __init__(self, data: bytes):
        self.b = _Buf(data)
        self._objects = {}
        self._gheap = {}
        b = self.b
        base = 0
        while True:   # the superblock sits at 0, 512, 1024, 2048, ...
            if data[base:base + 8] == b"\x89HDF\r\n\x1a\n":
                break
            base = 512 if base == 0 else base * 2
            if base >= len(data):
                raise FormatError("no HDF5 superblock")
        version = b.u(base + 8, 1)
        if version in (0, 1):
            b.so, b.sl = b.u(base + 13, 1), b.u(base + 14, 1)
            p = base + 24 + (4 if version == 1 else 0)
            p += 4 * b.so                       # base, free-space, end-of-file, driver addresses
            self.root = b.off(p + b.so)         # root symbol-table entry: object header address
        elif version in (2, 3):
            b.so, b.sl = b.u(base + 9, 1), b.u(base + 10, 1)
            p = base + 12
            self.root = b.off(p + 3 * b.so)     # base, extension, end-of-file, root object header
        else:
            raise FormatError(f"superblock version {version}")

    def obj(self, addr: int) -> _Object:
        if addr not in self._objects:
            self._objects[addr] = _Object(self, addr)
        return self._objects[addr]

    # -- groups ---------------------------------------------------------------------------------

    def links(self, addr: int) -> dict:
        """Name -> object header address of the group's hard links."""
        o = self.obj(addr)
        out = {}
        for _, _, p, _ in o.find(0x11):          # symbol table: B-tree v1 + local heap
            btree, heap = self.b.off(p), self.b.off(p + self.b.so)
            heap_data = self._local_heap(heap)
            self._walk_group_btree(btree, heap_data, out)
        for _, _, p, _ in o.find(0x06):          # link messages (compact new-style group)
            name, target = self._link(self.b, p)
            if target is not None:
                out[name] = target
        for _, _, p, _ in o.find(0x02):          # link info: links in a fractal heap (dense)
            flags = self.b.u(p + 1, 1)
            q = p + 2 + (8 if flags & 1 else 0)
            heap, names = self.b.off(q), self.b.off(q + self.b.so)
            if heap == _UNDEF:
                continue
            fh = _FractalHeap(self, heap)
            for rec in self._btree2_records(names):
                obj = fh.get(rec[4:4 + fh.id_len])     # type-5 record: name hash, heap ID
                name, target = self._link(_Buf(obj, self.b.so, self.b.sl), 0)
                if target is not None:
                    out[name] = target
        return out

    def _local_heap(self, addr: int) -> int:
        self.b.sig(addr, b"HEAP")
        return self.b.off(addr + 8 + 2 * self.b.sl)

    def _walk_group_btree(self, addr: int, heap_data: int, out: dict):
        b = self.b
        b.sig(addr, b"TREE")
        if b.u(addr + 4, 1) != 0:
            raise FormatError("group B-tree of the wrong node type")
        level, used = b.u(addr + 5, 1), b.u(addr + 6, 2)
        p = addr + 8 + 2 * b.so + b.sl        # past the siblings and key 0
        for _ in range(used):
            child = b.off(p)
            p += b.so + b.sl
            if level > 0:
                self._walk_group_btree(child, heap_data, out)
            else:
                b.sig(child, b"SNOD")
                n = b.u(child + 6, 2)
                e = child + 8
                for _ in range(n):
                    name = b.cstr(heap_data + b.off(e))
                    out[name] = b.off(e + b.so)
                    e += 2 * b.so + 24

    @staticmethod
    def _link(b: _Buf, p: int):
        flags = b.u(p + 1, 1)
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = b.u(q, 1)
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        n = 1 << (flags & 3)
        nlen = b.u(q, n)
        q += n
        name = b.d[q:q + nlen].decode("utf-8")
        q += nlen
        return name, (b.off(q) if ltype == 0 else None)

    def _btree2_records(self, addr: int) -> list:
        """All records of a version-2 B-tree (BTHD), in key order."""
        b = self.b
        b.sig(addr, b"BTHD")
        node_size, rec_size = b.u(addr + 6, 4), b.u(addr + 10, 2)
        depth = b.u(addr + 12, 2)
        root, root_n = b.off(addr + 16), b.u(addr + 16 + b.so, 2)
        # Field widths of the child pointers, per node level (HDF5 spec, "Version 2 B-trees").
        nmax, tmax, widths = 0, 0, []
        for level in range(depth + 1):
            if level == 0:
                w = (0, 0, 0)
            elif level == 1:
                w = (b.so, _nbytes(nmax), 0)
            else:
                w = (b.so, _nbytes(nmax), _nbytes(tmax))
            widths.append(w)
            ptr = sum(w)
            nmax = (node_size - 10 - ptr) // (rec_size + ptr)
            tmax = tmax * nmax if w[2] else nmax
        out = []

        def walk(node: int, nrec: int, level: int):
            sig = b"BTLF" if level == 0 else b"BTIN"
            b.sig(node, sig)
            p = node + 6
            recs = [bytes(b.d[p + i * rec_size:p + (i + 1) * rec_size]) for i in range(nrec)]
            if level == 0:
                out.extend(recs)
                return
            q = p + nrec * rec_size
            so, n1, n2 = widths[level]
            for i in range(nrec + 1):
                child, cn = b.off(q), b.u(q + so, n1)
                q += so + n1 + n2
                walk(child, cn, level - 1)
                if i < nrec:
                    out.append(recs[i])

        if root != _UNDEF and root_n:
            walk(root, root_n, depth)
        return out

    # -- datasets -------------------------------------------------------------------------------

    def _dataspace(self, p: int, b: _Buf = None):
        b = self.b if b is None else b
        version, ndim, flags = b.u(p, 1), b.u(p + 1, 1), b.u(p + 2, 1)
        if version == 1:
            q = p + 8
        elif version == 2:
            if b.u(p + 3, 1) == 2:
                return None          # null dataspace
            q = p + 4
        else:
            raise FormatError(f"dataspace version {version}")
        return tuple(b.length(q + i * b.sl) for i in range(ndim))

    def _filters(self, o: _Object):
        b = self.b
        out = []
        for _, _, p, _ in o.find(0x0B):
            version, nfilt = b.u(p, 1), b.u(p + 1, 1)
            q = p + (8 if version == 1 else 2)
            for _ in range(nfilt):
                fid = b.u(q, 2)
                if version == 1 or fid >= 256:
                    nlen = b.u(q + 2, 2)
                    q += 4
                else:
                    nlen = 0
                    q += 2
                nval = b.u(q + 2, 2)
                q += 4
                q += ((nlen + 7) // 8 * 8) if version == 1 else nlen
                vals = [b.u(q + 4 * i, 4) for i in range(nval)]
                q += 4 * nval + (4 if version == 1 and nval % 2 else 0)
                out.append((fid, vals))
        return out

    def _unfilter(self, raw: bytes, filters, mask: int, elem: int) -> bytes:
        for i in range(len(filters) - 1, -1, -1):
            if mask & (1 << i):
                continue
            fid, vals = filters[i]
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                size = vals[0] if vals else elem
                n = len(raw) // size
                a = np.frombuffer(raw[:n * size], np.uint8).reshape(size, n)
                raw = a.T.tobytes() + raw[n * size:]
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise FormatError(f"HDF5 filter {fid} is not supported")
        return raw

    def read_dataset(self, addr: int):
        """(shape, datatype, C-order array) of the dataset at `addr`."""
        o = self.obj(addr)
        b = self.b
        ds = o.find(0x01)
        dt = o.find(0x03)
        if not ds or not dt:
            raise FormatError("object is not a dataset")
        shape = self._dataspace(ds[0][2])
        dtype = _Datatype(b, dt[0][2])
        if dtype.dtype is None:
            raise FormatError(f"dataset of datatype class {dtype.cls}")
        if shape is None:
            return (), dtype, np.zeros((), dtype.dtype)
        n = int(np.prod(shape, dtype=np.int64))
        out = bytearray(self._fill(o, dtype) * n)
        lay = o.find(0x08)[0][2]
        version, cls = b.u(lay, 1), b.u(lay + 1, 1)
        if version < 3:
            raise FormatError(f"data layout version {version}")
        esize = dtype.size
        if cls == 0:                              # compact
            size = b.u(lay + 2, 2)
            out[:size] = b.d[lay + 4:lay + 4 + size]
        elif cls == 1:                            # contiguous
            a, size = b.off(lay + 2), b.length(lay + 2 + b.so)
            if a != _UNDEF:
                out[:size] = b.d[a:a + size]
        elif cls == 2:
            filters = self._filters(o)
            if version == 3:
                dimensionality = b.u(lay + 2, 1)
                btree = b.off(lay + 3)
                chunk = tuple(b.u(lay + 3 + b.so + 4 * i, 4) for i in range(dimensionality - 1))
                chunks = [] if btree == _UNDEF else self._chunk_btree(btree, len(chunk))
            else:
                chunk, chunks = self._layout4(lay, shape, esize)
            for offs, size, mask, caddr in chunks:
                raw = self._unfilter(bytes(b.d[caddr:caddr + size]), filters, mask, esize)
                self._place(out, shape, chunk, offs, raw, esize)
        else:
            raise FormatError(f"layout class {cls}")
        arr = np.frombuffer(bytes(out), dtype.dtype, count=n).reshape(shape)
        return shape, dtype, arr

    def _fill(self, o: _Object, dtype: _Datatype) -> bytes:
        b = self.b
        for _, _, p, size in o.find(0x05):
            version = b.u(p, 1)
            if version in (1, 2):
                defined = b.u(p + 3, 1)
                if defined and size >= 8 + dtype.size:
                    if b.u(p + 4, 4) == dtype.size:
                        return bytes(b.d[p + 8:p + 8 + dtype.size])
            elif version == 3:
                flags = b.u(p + 1, 1)
                if flags & 0x20 and b.u(p + 2, 4) == dtype.size:
                    return bytes(b.d[p + 6:p + 6 + dtype.size])
        return bytes(dtype.size)

    def _chunk_btree(self, addr: int, rank: int):
        b = self.b
        b.sig(addr, b"TREE")
        if b.u(addr + 4, 1) != 1:
            raise FormatError("chunk B-tree of the wrong node type")
        level, used = b.u(addr + 5, 1), b.u(addr + 6, 2)
        ksize = 8 + 8 * (rank + 1)
        p = addr + 8 + 2 * b.so
        out = []
        for _ in range(used):
            size, mask = b.u(p, 4), b.u(p + 4, 4)
            offs = tuple(b.u(p + 8 + 8 * i, 8) for i in range(rank))
            child = b.off(p + ksize)
            if level > 0:
                out.extend(self._chunk_btree(child, rank))
            else:
                out.append((offs, size, mask, child))
            p += ksize + b.so
        return out

    def _layout4(self, lay: int, shape: tuple, esize: int):
        b = self.b
        rank = len(shape)
        flags = b.u(lay + 2, 1)
        dimensionality = b.u(lay + 3, 1)
        enc = b.u(lay + 4, 1)
        q = lay + 5
        dims = [b.u(q + enc * i, enc) for i in range(dimensionality)]
        q += enc * dimensionality
        chunk = tuple(dims[:rank])
        index = b.u(q, 1)
        q += 1
        nchunks = [0] * rank
        grid = []
        for i in range(rank):
            nchunks[i] = -(-shape[i] // chunk[i])
        for flat in range(int(np.prod(nchunks))):
            idx, rem = [], flat
            for i in range(rank - 1, -1, -1):
                idx.append(rem % nchunks[i])
                rem //= nchunks[i]
            grid.append(tuple(c * s for c, s in zip(reversed(idx), chunk)))
        full = int(np.prod(chunk)) * esize
        if index == 1:                         # single chunk
            if flags & 0x02:
                size, mask = b.length(q), b.u(q + b.sl, 4)
                addr = b.off(q + b.sl + 4)
            else:
                size, mask, addr = full, 0, b.off(q)
            return chunk, [] if addr == _UNDEF else [((0,) * rank, size, mask, addr)]
        if index == 2:                         # implicit: chunks stored contiguously in order
            addr = b.off(q)
            return chunk, [(g, full, 0, addr + k * full) for k, g in enumerate(grid)]
        if index == 3:                         # fixed array
            q += 1
            addr = b.off(q)
            b.sig(addr, b"FAHD")
            ctype = b.u(addr + 5, 1)
            esz = b.u(addr + 6, 1)
            page_bits = b.u(addr + 7, 1)
            if b.length(addr + 8) > (1 << page_bits):
                raise FormatError("paged fixed-array chunk indices are not read")
            dblk = b.off(addr + 8 + b.sl)
            b.sig(dblk, b"FADB")
            p = dblk + 6 + b.so
            npages = len(grid)
            out = []
            for k in range(npages):
                if ctype == 0:
                    caddr, size, mask = b.off(p), full, 0
                else:
                    caddr = b.off(p)
                    nb = esz - b.so - 4
                    size, mask = b.u(p + b.so, nb), b.u(p + b.so + nb, 4)
                p += esz
                if caddr != _UNDEF:
                    out.append((grid[k], size, mask, caddr))
            return chunk, out
        raise FormatError(f"chunk index type {index} is not supported")

    @staticmethod
    def _place(out: bytearray, shape, chunk, offs, raw: bytes, esize: int):
        cshape = tuple(chunk)
        c = np.frombuffer(raw[:int(np.prod(cshape)) * esize], np.uint8).reshape(cshape + (esize,))
        dst = np.frombuffer(out, np.uint8).reshape(tuple(shape) + (esize,))
        sl_dst, sl_src = [], []
        for o_, cs, s in zip(offs, cshape, shape):
            e = min(o_ + cs, s)
            sl_dst.append(slice(o_, e))
            sl_src.append(slice(0, e - o_))
        dst[tuple(sl_dst)] = c[tuple(sl_src)]

    # -- attributes -----------------------------------------------------------------------------

    def attributes(self, addr: int) -> dict:
        """Name -> (datatype, shape, buffer, value offset) of an object's attributes."""
        out = {}
        o = self.obj(addr)
        for _, _, p, _ in o.find(0x0C):           # compact: attribute messages in the header
            name, val = self._attribute(self.b, p)
            out[name] = val
        for _, _, p, _ in o.find(0x15):           # attribute info: dense storage
            b = self.b
            flags = b.u(p + 1, 1)
            q = p + 2 + (2 if flags & 1 else 0)
            heap, names = b.off(q), b.off(q + b.so)
            if heap == _UNDEF or names == _UNDEF:
                continue
            fh = _FractalHeap(self, heap)
            for rec in self._btree2_records(names):   # type-8 record: heap ID first
                obj = _Buf(fh.get(rec[:8]), b.so, b.sl)
                name, val = self._attribute(obj, 0)
                out[name] = val
        return out

    def _attribute(self, b: _Buf, p: int):
        version = b.u(p, 1)
        nsize, tsize, ssize = b.u(p + 2, 2), b.u(p + 4, 2), b.u(p + 6, 2)
        q = p + 8 + (1 if version == 3 else 0)
        pad = (lambda x: (x + 7) // 8 * 8) if version == 1 else (lambda x: x)
        name = b.d[q:q + nsize].split(b"\0")[0].decode("utf-8")
        q += pad(nsize)
        dtype = _Datatype(b, q)
        q += pad(tsize)
        shape = self._dataspace(q, b)
        q += pad(ssize)
        return name, (dtype, shape, b, q)

    def attribute_value(self, dtype: _Datatype, shape, b: _Buf, p: int):
        n = 1 if not shape else int(np.prod(shape))
        if dtype.dtype is not None:
            a = np.frombuffer(b.d, dtype.dtype, count=n, offset=p)
            if dtype.cls == 3:
                return b"".join(a.tolist()).split(b"\0")[0].decode("utf-8", "replace")
            return a.reshape(shape) if shape else a[0]
        if dtype.cls == 9:
            vals = []
            for i in range(n):
                q = p + i * (4 + b.so + 4)
                length, coll, idx = b.u(q, 4), b.off(q + 4), b.u(q + 4 + b.so, 4)
                data = self._global_object(coll, idx) if length else b""
                if dtype.vlen_string:
                    vals.append(data.decode("utf-8", "replace"))
                elif dtype.base.cls == 7:
                    vals.append([int.from_bytes(data[k:k + b.so], "little") for k in range(0, len(data), b.so)])
                else:
                    vals.append(data)
            return vals
        return None

    def _global_object(self, coll: int, idx: int) -> bytes:
        if coll not in self._gheap:
            b = self.b
            b.sig(coll, b"GCOL")
            size = b.length(coll + 8)
            objs = {}
            p, end = coll + 8 + b.sl, coll + size
            while p + 8 + b.sl <= end:
                i = b.u(p, 2)
                if i == 0:
                    break
                osize = b.length(p + 8)
                objs[i] = bytes(b.d[p + 8 + b.sl:p + 8 + b.sl + osize])
                p += 8 + b.sl + (osize + 7) // 8 * 8
            self._gheap[coll] = objs
        return self._gheap[coll][idx]


# ---------------------------------------------------------------------------------------------
# The netCDF view.

class Variable:
    """A NetCDF variable: dimension names, shape and (lazily read) values in C order."""

    def __init__(self, name: str, dims: tuple, shape: tuple, loader, attrs: dict):
        self.name = name
        self.dimensions = dims
        self.shape = shape
        self._loader = loader
        self.attributes = attrs

    @property
    def data(self) -> np.ndarray:
        a = self._loader()
        if a.dtype.kind in "iuf" and a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        return a


class Dataset:
    """Dimensions and variables of a NetCDF file (root group), read-only."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            data = fh.read()
        self.path = path
        self.dimensions: dict = {}
        self.variables: dict = {}
        if data[:8] == b"\x89HDF\r\n\x1a\n" or b"\x89HDF\r\n\x1a\n" in data[:4096:512]:
            self._open_hdf5(data)
        elif data[:3] == b"CDF":
            self._open_classic(path)
        else:
            raise FormatError(f"{path}: neither NetCDF-4 (HDF5) nor NetCDF-3")

    def _open_classic(self, path: str):
        h = _read_classic_header(path)
        self.dimensions.update({name: size for name, size in h["dims"]})
        for v in h["vars"]:
            self.variables[v["name"]] = Variable(v["name"], v["dims"], v["shape"],
                                                 (lambda v=v: _read_classic_var(path, h, v)), v["attrs"])

    def _open_hdf5(self, data: bytes):
        h = _Hdf5(data)
        links = h.links(h.root)
        infos = {}
        scales = {}
        for name, addr in links.items():
            o = h.obj(addr)
            if not o.find(0x08):       # not a dataset (a subgroup or a named datatype)
                continue
            attrs = h.attributes(addr)
            shape = h._dataspace(o.find(0x01)[0][2]) or ()
            infos[name] = (addr, shape, attrs)
            cls = attrs.get("CLASS")
            if cls is not None and h.attribute_value(*cls) == "DIMENSION_SCALE":
                scales[addr] = name
                self.dimensions[name] = int(shape[0]) if shape else 0
        for name, (addr, shape, attrs) in infos.items():
            dims = ()
            if addr in scales:
                dims = (name,)
                label = attrs.get("NAME")
                is_var = label is None or not h.attribute_value(*label).startswith(
                    "This is a netCDF dimension but not a netCDF variable")
                if not is_var:
                    continue
            elif "DIMENSION_LIST" in attrs:
                refs = h.attribute_value(*attrs["DIMENSION_LIST"])
                dims = tuple(scales.get(r[0], f"dim{i}") if r else f"dim{i}" for i, r in enumerate(refs))
            elif shape:
                dims = tuple(f"phony_dim_{i}" for i in range(len(shape)))
            user = {}
            for k, v in attrs.items():
                if k in ("CLASS", "NAME", "DIMENSION_LIST", "REFERENCE_LIST") or k.startswith("_Netcdf4"):
                    continue
                try:
                    user[k] = h.attribute_value(*v)
                except FormatError:
                    pass

            def load(a=addr):
                return h.read_dataset(a)[2]
            self.variables[name] = Variable(name, dims, tuple(shape), load, user)

    def get_dimension_size(self, name: str) -> int:
        if name not in self.dimensions:
            raise KeyError(f"{self.path}: no dimension {name!r}")
        return self.dimensions[name]

    def get_variable(self, name: str, dims=None) -> np.ndarray:
        """Values of `name` in C order; with `dims`, checks the shape as the reference does."""
        if name not in self.variables:
            raise KeyError(f"{self.path}: no variable {name!r}")
        v = self.variables[name]
        a = v.data
        if dims is not None and tuple(dims) != tuple(a.shape):
            if int(np.prod(dims)) != a.size:
                raise ValueError(f"{name}: shape {a.shape} does not match the requested {tuple(dims)}")
            a = a.reshape(tuple(dims))
        return a

    def get_strings(self, name: str, trim: bool = True) -> list:
        """A char variable of shape (n, string_len) as n strings (Radiation_solver.cpp:48-68)."""
        a = self.get_variable(name)
        rows = a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(1, -1)
        out = []
        for r in rows:
            s = b"".join(r.tolist()).split(b"\0")[0].decode("utf-8")
            out.append(s.strip() if trim else s)
        return out


# ---------------------------------------------------------------------------------------------
# NetCDF-3 (classic, 64-bit offset, 64-bit data): big-endian header of dimension, attribute and
# variable lists, then the fixed-size variables in order (4-byte aligned) and the records.

_NC_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8", 7: ">u1", 8: ">u2", 9: ">u4",
             10: ">i8", 11: ">u8"}
_NC_CODES = {"i1": 1, "S1": 2, "i2": 3, "i4": 4, "f4": 5, "f8": 6, "u1": 7, "u2": 8, "u4": 9, "i8": 10,
             "u8": 11}
_TAG_DIM, _TAG_VAR, _TAG_ATT = 10, 11, 12


def _pad4(n: int) -> int:
    return (n + 3) & ~3


def _read_classic_header(path: str) -> dict:
    with open(path, "rb") as fh:
        head = fh.read(4)
        if head[:3] != b"CDF" or head[3] not in (1, 2, 5):
            raise FormatError(f"{path}: not a NetCDF-3 file")
        version = head[3]
        # The header ends before the first variable's data; read generously and grow on demand.
        size = 1 << 16
        while True:
            fh.seek(0)
            buf = fh.read(size)
            try:
                return _parse_classic_header(buf, version, path)
            except IndexError:
                if len(buf) < size:
                    raise FormatError(f"{path}: truncated NetCDF-3 header")
                size *= 4


def _parse_classic_header(buf: bytes, version: int, path: str) -> dict:
    pos = 4
    big = version == 5

    def take(n):
        nonlocal pos
        if pos + n > len(buf):
            raise IndexError
        b = buf[pos:pos + n]
        pos += n
        return b

    def i32():
        return struct.unpack(">i", take(4))[0]

    def nonneg():
        return struct.unpack(">q", take(8))[0] if big else struct.unpack(">i", take(4))[0]

    def offset():
        return struct.unpack(">q", take(8))[0] if version in (2, 5) else struct.unpack(">i", take(4))[0]

    def name():
        n = nonneg()
        return take(_pad4(n))[:n].decode("utf-8")

    def atts():
        tag, n = i32(), nonneg()
        if tag == 0 and n == 0:
            return {}
        if tag != _TAG_ATT:
            raise FormatError(f"{path}: bad attribute list tag {tag}")
        out = {}
        for _ in range(n):
            k = name()
            t = i32()
            cnt = nonneg()
            dt = np.dtype(_NC_TYPES[t])
            raw = take(_pad4(cnt * dt.itemsize))[:cnt * dt.itemsize]
            if t == 2:
                out[k] = raw.split(b"\0")[0].decode("utf-8", "replace")
            else:
                a = np.frombuffer(raw, dt).astype(dt.newbyteorder("="))
                out[k] = a[0] if a.size == 1 else a
        return out

    numrecs = nonneg()
    dims = []
    tag, n = i32(), nonneg()
    if not (tag == 0 and n == 0):
        if tag != _TAG_DIM:
            raise FormatError(f"{path}: bad dimension list tag {tag}")
        for _ in range(n):
            dims.append((name(), nonneg()))
    gatts = atts()
    tag, n = i32(), nonneg()
    vars_ = []
    if not (tag == 0 and n == 0):
        if tag != _TAG_VAR:
            raise FormatError(f"{path}: bad variable list tag {tag}")
        for _ in range(n):
            vname = name()
            nd = nonneg()
            ids = [nonneg() for _ in range(nd)]
            vatts = atts()
            t = i32()
            vsize = nonneg()
            begin = offset()
            vdims = tuple(dims[i][0] for i in ids)
            rec = bool(ids) and dims[ids[0]][1] == 0
            shape = tuple((numrecs if (rec and j == 0) else dims[i][1]) for j, i in enumerate(ids))
            vars_.append(dict(name=vname, dims=vdims, shape=shape, type=t, vsize=vsize, begin=begin,
                              record=rec, attrs=vatts))
    recsize = sum(v["vsize"] for v in vars_ if v["record"])
    if sum(1 for v in vars_ if v["record"]) == 1:
        v = next(v for v in vars_ if v["record"])
        recsize = int(np.prod(v["shape"][1:], dtype=np.int64)) * np.dtype(_NC_TYPES[v["type"]]).itemsize
    return dict(version=version, numrecs=numrecs, dims=dims, gatts=gatts, vars=vars_, recsize=recsize)


def _read_classic_var(path: str, h: dict, v: dict) -> np.ndarray:
    dt = np.dtype(_NC_TYPES[v["type"]])
    with open(path, "rb") as fh:
        if not v["record"]:
            count = int(np.prod(v["shape"], dtype=np.int64))
            fh.seek(v["begin"])
            raw = fh.read(count * dt.itemsize)
            if len(raw) != count * dt.itemsize:
                raise FormatError(f"{path}: variable {v['name']} runs past the end of the file")
            a = np.frombuffer(raw, dt).reshape(v["shape"])
        else:
            per = int(np.prod(v["shape"][1:], dtype=np.int64)) * dt.itemsize
            parts = []
            for r in range(h["numrecs"]):
                fh.seek(v["begin"] + r * h["recsize"])
                parts.append(fh.read(per))
            a = np.frombuffer(b"".join(parts), dt).reshape(v["shape"])
    if dt.kind in "iuf":
        a = a.astype(dt.newbyteorder("="))
    return a.copy()


class Writer:
    """A NetCDF-3 file written in one go: add dimensions and variables, then close().

    Mirrors the parts of Netcdf_file / Netcdf_variable the driver uses (add_dimension,
    add_variable, insert of the whole variable); variables are C-ordered numpy arrays whose shape
    matches their dimensions.  The file is CDF-2 (64-bit offset) unless a fixed-size variable other
    than the last passes 4 GiB, then CDF-5."""

    def __init__(self, path: str, version: int | None = None):
        if version not in (None, 2, 5):
            raise ValueError("NetCDF-3 version must be 2 (64-bit offset) or 5 (64-bit data)")
        self.path = path
        self.version = version
        self.dims: list = []
        self.vars: list = []
        self.gatts: dict = {}

    def add_dimension(self, name: str, size: int):
        if any(n == name for n, _ in self.dims):
            if dict(self.dims)[name] != size:
                raise ValueError(f"dimension {name!r} exists with another size")
            return
        if size <= 0:
            raise ValueError(f"dimension {name!r}: size {size} (record dimensions are not written)")
        self.dims.append((name, int(size)))

    def add_variable(self, name: str, dims: tuple, data, attrs: dict | None = None):
        d = dict(self.dims)
        for n in dims:
            if n not in d:
                raise KeyError(f"variable {name!r}: no dimension {n!r}")
        a = np.asarray(data)
        if a.dtype.kind == "f" and a.dtype.itemsize not in (4, 8):
            a = a.astype(np.float64)
        if a.dtype == np.bool_:
            a = a.astype(np.int8)
        shape = tuple(d[n] for n in dims)
        if a.size != int(np.prod(shape, dtype=np.int64)):
            raise ValueError(f"variable {name!r}: {a.size} values for dimensions {dims} {shape}")
        code = a.dtype.kind + str(a.dtype.itemsize) if a.dtype.kind != "S" else "S1"
        if code not in _NC_CODES:
            raise TypeError(f"variable {name!r}: type {a.dtype} has no NetCDF-3 counterpart")
        self.vars.append((name, tuple(dims), np.ascontiguousarray(a.reshape(shape)), attrs or {}))

    def close(self):
        big = any(v[2].nbytes >= (1 << 32) - 4 for v in self.vars[:-1])
        version = self.version or (5 if big else 2)
        if version == 2 and big:
            raise ValueError("a fixed-size variable other than the last passes CDF-2's 4 GiB")
        dim_index = {n: i for i, (n, _) in enumerate(self.dims)}

        def nonneg(x):
            return struct.pack(">q", x) if version == 5 else struct.pack(">i", x)

        def name(s):
            b = s.encode("utf-8")
            return nonneg(len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))

        def atts(d):
            if not d:
                return struct.pack(">i", 0) + nonneg(0)
            out = struct.pack(">i", _TAG_ATT) + nonneg(len(d))
            for k, val in d.items():
                if isinstance(val, str):
                    raw, t, n = val.encode("utf-8"), 2, len(val.encode("utf-8"))
                else:
                    a = np.atleast_1d(np.asarray(val))
                    code = a.dtype.kind + str(a.dtype.itemsize)
                    t = _NC_CODES[code]
                    raw, n = a.astype(np.dtype(_NC_TYPES[t])).tobytes(), a.size
                out += name(k) + struct.pack(">i", t) + nonneg(n) + raw + b"\0" * (_pad4(len(raw)) - len(raw))
            return out

        def header(begins):
            h = b"CDF" + bytes([version]) + nonneg(0)
            if self.dims:
                h += struct.pack(">i", _TAG_DIM) + nonneg(len(self.dims))
                for n, size in self.dims:
                    h += name(n) + nonneg(size)
            else:
                h += struct.pack(">i", 0) + nonneg(0)
            h += atts(self.gatts)
            if self.vars:
                h += struct.pack(">i", _TAG_VAR) + nonneg(len(self.vars))
                for (vn, vd, a, va), b in zip(self.vars, begins):
                    code = a.dtype.kind + str(a.dtype.itemsize) if a.dtype.kind != "S" else "S1"
                    vsize = _pad4(a.nbytes)
                    if version == 2 and vsize >= (1 << 32):
                        vsize = (1 << 32) - 1          # the last variable of a CDF-2 file
                    h += name(vn) + nonneg(len(vd)) + b"".join(nonneg(dim_index[x]) for x in vd)
                    h += atts(va) + struct.pack(">i", _NC_CODES[code]) + nonneg(vsize) + struct.pack(">q", b)
            else:
                h += struct.pack(">i", 0) + nonneg(0)
            return h

        hlen = len(header([0] * len(self.vars)))
        begins, off = [], hlen
        for _, _, a, _ in self.vars:
            begins.append(off)
            off += _pad4(a.nbytes)
        h = header(begins)
        assert len(h) == hlen
        with open(self.path, "wb") as fh:
            fh.write(h)
            for _, _, a, _ in self.vars:
                if a.dtype.kind == "S":
                    raw = a.astype("S1").tobytes()
                else:
                    raw = a.astype(a.dtype.newbyteorder(">")).tobytes()
                fh.write(raw)
                fh.write(b"\0" * (_pad4(len(raw)) - len(raw)))


def write_schema(path: str, schema, version: int | None = None):
    """A coefficients.Schema (dimensions and C-ordered variables) as a NetCDF-3 file."""
    w = Writer(path, version)
    for name, size in schema.dims.items():
        w.add_dimension(name, size)
    for name, (dims, a) in schema.vars.items():
        w.add_variable(name, dims, a)
    w.close()


def to_schema(ds: Dataset):
    """All dimensions and variables of a coefficient file as a bundle schema."""
    from .coefficients import Schema
    s = Schema()
    for name, size in ds.dimensions.items():
        s.dim(name, size)
    for name, v in ds.variables.items():
        a = v.data
        if a.dtype.kind == "S" and a.dtype.itemsize != 1:
            a = a.view("S1").reshape(a.shape + (a.dtype.itemsize,))
        if a.dtype.kind == "f" and a.dtype.itemsize == 2:
            a = a.astype("<f4")
        if a.dtype.kind in "iu" and a.dtype.itemsize != 4 and a.dtype.itemsize != 1:
            a = a.astype("<i4")
        if a.dtype.kind == "u" and a.dtype.itemsize == 1:
            a = a.astype("i1")
        if a.dtype.kind == "u" and a.dtype.itemsize == 4:
            a = a.astype("<i4")
        dims = tuple(v.dimensions)
        if a.ndim == len(dims) + 1:         # fixed-length strings split into characters
            dims = dims + ("string_len",)
        s.put(name, dims, a)
    return s


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("usage: python -m rrtmgp_amd.netcdf <file.nc> <bundle_dir>", file=sys.stderr)
        return 2
    from .coefficients import write_bundle
    write_bundle(argv[1], to_schema(Dataset(argv[0])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
