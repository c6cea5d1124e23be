"""Classic netCDF (CDF-1 / CDF-2) files without the netCDF library.

The reference writes its netCDF snapshots through netCDF-Fortran with
``nf90_create(..., NF90_CLOBBER)`` (src/swpc_3d/m_snap.f90:623-741,
942-971, 2185-2344) and reads random-medium volumes the same way
(src/shared/m_rdrmed.f90:24-134).  Those are classic-format files, whose
layout is the published netCDF classic format specification: a big-endian
header (magic ``CDF\\x01`` or ``CDF\\x02``, the record count, the dimension,
global-attribute and variable lists, each variable carrying its type, its
per-record size ``vsize`` and its file offset ``begin``), then the fixed-size
variables in definition order, then the records, each holding one slab of
every record variable in definition order.

``Writer`` builds a file in define mode, fixes the layout at ``enddef``
(fixed-size variables pre-filled with the type's default fill value, as the
library's default fill mode does) and then takes whole-variable and
per-record writes; records skipped over are filled.  An attribute may be
changed after ``enddef`` when its type and length stay the same (the
header is rewritten in place), which covers the reference's
``actual_range`` updates.  ``read`` parses either version.
"""
from __future__ import annotations

import os
import struct

import numpy as np

NC_BYTE, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE = 1, 2, 3, 4, 5, 6
_DIM, _VAR, _ATT = 0x0A, 0x0B, 0x0C

_NP = {NC_BYTE: ">i1", NC_CHAR: "S1", NC_SHORT: ">i2", NC_INT: ">i4", NC_FLOAT: ">f4", NC_DOUBLE: ">f8"}
_SIZE = {NC_BYTE: 1, NC_CHAR: 1, NC_SHORT: 2, NC_INT: 4, NC_FLOAT: 4, NC_DOUBLE: 8}
_BY_NAME = {"i1": NC_BYTE, "c": NC_CHAR, "i2": NC_SHORT, "i4": NC_INT, "f4": NC_FLOAT, "f8": NC_DOUBLE}
# default fill values of the classic format
FILL = {NC_BYTE: -127, NC_CHAR: 0, NC_SHORT: -32767, NC_INT: -2147483647,
        NC_FLOAT: 9.9692099683868690e36, NC_DOUBLE: 9.9692099683868690e36}


class CDFError(ValueError):
    pass


def _pad4(n):
    return (n + 3) & ~3


def _name(s):
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _att_value(v):
    """(nc_type, big-endian payload, nelems) of an attribute value: str -> char,
    numpy float32 -> float, numpy float64 / Python float -> double, integers
    -> int; a list or array takes the type of its elements."""
    if isinstance(v, (str, bytes)):
        b = v.encode() if isinstance(v, str) else v
        return NC_CHAR, b, len(b)
    a = np.asarray(v)
    if a.dtype == np.float32:
        t = NC_FLOAT
    elif a.dtype.kind == "f":
        t = NC_DOUBLE
    elif a.dtype.kind in "iu":
        t = NC_INT if a.dtype.itemsize >= 4 or a.dtype.kind == "u" else (NC_SHORT if a.dtype.itemsize == 2 else NC_BYTE)
    else:
        raise CDFError(f"unsupported attribute value {v!r}")
    a = a.reshape(-1).astype(_NP[t])
    return t, a.tobytes(), a.size


def _attrs(atts):
    if not atts:
        return struct.pack(">ii", 0, 0)
    out = [struct.pack(">ii", _ATT, len(atts))]
    for k, v in atts.items():
        t, payload, n = _att_value(v)
        out.append(_name(k) + struct.pack(">ii", t, n) + payload + b"\0" * (_pad4(len(payload)) - len(payload)))
    return b"".join(out)


class _Var:
    def __init__(self, name, nctype, dims, attrs):
        self.name, self.nctype, self.dims, self.attrs = name, nctype, tuple(dims), dict(attrs)
        self.begin = 0
        self.vsize = 0
        self.shape = ()
        self.record = False


class Writer:
    """Classic-format netCDF writer (version 1: 32-bit offsets, 2: 64-bit)."""

    def __init__(self, path, version=1):
        if version not in (1, 2):
            raise CDFError("version must be 1 or 2")
        self.path, self.version = path, version
        self.dims = []          # (name, size or None for the record dimension)
        self.gatts = {}
        self.vars = []
        self.numrecs = 0
        self.defining = True
        self.f = None

    # define mode -----------------------------------------------------------
    def dim(self, name, size):
        if not self.defining:
            raise CDFError("dimensions are defined before enddef")
        if size is None and any(s is None for _, s in self.dims):
            raise CDFError("only one record dimension")
        self.dims.append((name, size))
        return len(self.dims) - 1

    def var(self, name, kind, dims, attrs=None):
        if not self.defining:
            raise CDFError("variables are defined before enddef")
        ids = [self._dimid(d) for d in dims]
        for n, i in enumerate(ids):
            if self.dims[i][1] is None and n != 0:
                raise CDFError("the record dimension must be the slowest")
        v = _Var(name, _BY_NAME[kind], ids, attrs or {})
        self.vars.append(v)
        return v

    def gatt(self, name, value):
        self.gatts[name] = value
        if not self.defining:
            self._rewrite_header_check()

    def att(self, var, name, value):
        v = self._var(var)
        v.attrs[name] = value
        if not self.defining:
            self._rewrite_header_check()

    def _dimid(self, d):
        if isinstance(d, int):
            return d
        for i, (n, _) in enumerate(self.dims):
            if n == d:
                return i
        raise CDFError(f"unknown dimension {d!r}")

    def _var(self, v):
        if isinstance(v, _Var):
            return v
        for x in self.vars:
            if x.name == v:
                return x
        raise CDFError(f"unknown variable {v!r}")

    # layout ----------------------------------------------------------------
    def _header(self):
        h = [b"CDF" + bytes([self.version]), struct.pack(">i", self.numrecs)]
        if self.dims:
            h.append(struct.pack(">ii", _DIM, len(self.dims)))
            for n, s in self.dims:
                h.append(_name(n) + struct.pack(">i", 0 if s is None else s))
        else:
            h.append(struct.pack(">ii", 0, 0))
        h.append(_attrs(self.gatts))
        if self.vars:
            h.append(struct.pack(">ii", _VAR, len(self.vars)))
            for v in self.vars:
                h.append(_name(v.name) + struct.pack(">i", len(v.dims)) +
                         b"".join(struct.pack(">i", d) for d in v.dims) + _attrs(v.attrs) +
                         struct.pack(">ii", v.nctype, v.vsize) +
                         (struct.pack(">i", v.begin) if self.version == 1 else struct.pack(">q", v.begin)))
        else:
            h.append(struct.pack(">ii", 0, 0))
        return b"".join(h)

    def enddef(self):
        if not self.defining:
            return
        for v in self.vars:
            shp = tuple(self.dims[d][1] for d in v.dims)
            v.record = bool(shp) and shp[0] is None
            v.shape = shp[1:] if v.record else shp
            v.vsize = _pad4(int(np.prod(v.shape, dtype=np.int64)) * _SIZE[v.nctype])
        hsize = len(self._header())
        off = _pad4(hsize)
        for v in self.vars:
            if not v.record:
                v.begin = off
                off += v.vsize
        self.begin_rec = off
        self.recsize = 0
        for v in self.vars:
            if v.record:
                v.begin = off + self.recsize
                self.recsize += v.vsize
        recs = [v for v in self.vars if v.record]
        self.single_rec = len(recs) == 1
        if self.single_rec:   # a single record variable is not padded between records
            self.recsize = int(np.prod(recs[0].shape, dtype=np.int64)) * _SIZE[recs[0].nctype]
        limit = 2**31 - 1 if self.version == 1 else 2**63 - 1
        if self.begin_rec > limit:
            raise CDFError("fixed-size variables exceed the offset range of this version")
        self.hsize = hsize
        self.f = open(self.path, "wb+")
        self.f.write(self._header())
        self.f.write(b"\0" * (_pad4(hsize) - hsize))
        for v in self.vars:
            if not v.record:
                self.f.seek(v.begin)
                self.f.write(self._fill(v, int(np.prod(v.shape, dtype=np.int64))))
        self.defining = False

    def _fill(self, v, n, pad=True):
        b = np.full(n, FILL[v.nctype], dtype=_NP[v.nctype]).tobytes() if v.nctype != NC_CHAR else b"\0" * n
        return b + b"\0" * (_pad4(len(b)) - len(b)) if pad else b

    def _rewrite_header_check(self):
        h = self._header()
        if len(h) != self.hsize:
            raise CDFError("after enddef an attribute may change its value only, not its type or length")
        self.f.seek(0)
        self.f.write(h)

    # data ------------------------------------------------------------------
    def put(self, var, data):
        v = self._var(var)
        if self.defining:
            self.enddef()
        if v.record:
            raise CDFError(f"{v.name} is a record variable: use put_record")
        a = np.ascontiguousarray(data, dtype=_NP[v.nctype])
        if a.size != int(np.prod(v.shape, dtype=np.int64)):
            raise CDFError(f"{v.name}: {a.size} values for shape {v.shape}")
        self.f.seek(v.begin)
        self.f.write(a.tobytes())

    def put_record(self, var, rec, data):
        """Write record rec (0-based) of a record variable; records between
        the current count and rec are filled first."""
        v = self._var(var)
        if self.defining:
            self.enddef()
        if not v.record:
            raise CDFError(f"{v.name} is not a record variable")
        a = np.ascontiguousarray(data, dtype=_NP[v.nctype])
        if a.size != int(np.prod(v.shape, dtype=np.int64)):
            raise CDFError(f"{v.name}: {a.size} values for a record of shape {v.shape}")
        while self.numrecs <= rec:
            self._fill_record(self.numrecs)
            self.numrecs += 1
        self.f.seek(v.begin + rec * self.recsize)
        self.f.write(a.tobytes())

    def _fill_record(self, r):
        for v in self.vars:
            if v.record:
                self.f.seek(v.begin + r * self.recsize)
                self.f.write(self._fill(v, int(np.prod(v.shape, dtype=np.int64)), pad=not self.single_rec))

    def close(self):
        if self.defining:
            self.enddef()
        if self.f is None:
            return
        self.f.seek(0)
        self.f.write(self._header())
        self.f.close()
        self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# reader
# ---------------------------------------------------------------------------
class Dataset:
    """Parsed classic-format file: dims {name: size}, record dimension name,
    numrecs, gatts {name: value}, variables {name: Variable}."""

    def __init__(self, path):
        self.path = path
        with open(path, "rb") as f:
            self._b = f.read()
        b = self._b
        if b[:3] != b"CDF" or b[3] not in (1, 2):
            raise CDFError(f"{path}: not a classic netCDF file")
        self.version = b[3]
        self._p = 4
        self.numrecs = self._i()
        self.dim_list = []
        tag, n = self._i(), self._i()
        if tag not in (0, _DIM):
            raise CDFError("bad dimension list")
        for _ in range(n):
            nm = self._s()
            self.dim_list.append((nm, self._i()))
        self.dims = {nm: (self.numrecs if s == 0 else s) for nm, s in self.dim_list}
        self.record_dim = next((nm for nm, s in self.dim_list if s == 0), None)
        self.gatts = self._atts()
        tag, n = self._i(), self._i()
        if tag not in (0, _VAR):
            raise CDFError("bad variable list")
        vars_ = []
        for _ in range(n):
            nm = self._s()
            nd = self._i()
            ids = [self._i() for _ in range(nd)]
            atts = self._atts()
            t, vsize = self._i(), self._i()
            begin = self._i() if self.version == 1 else self._q()
            vars_.append((nm, ids, atts, t, vsize, begin))
        recs = [v for v in vars_ if v[1] and self.dim_list[v[1][0]][1] == 0]
        recsize = sum(v[4] for v in recs)
        if len(recs) == 1:
            shp = [self.dim_list[d][1] for d in recs[0][1][1:]]
            recsize = int(np.prod(shp, dtype=np.int64)) * _SIZE[recs[0][3]]
        self.variables = {}
        for nm, ids, atts, t, vsize, begin in vars_:
            shp = [self.dim_list[d][1] for d in ids]
            rec = bool(ids) and shp[0] == 0
            self.variables[nm] = Variable(self, nm, [self.dim_list[d][0] for d in ids], atts, t,
                                          tuple(shp[1:]) if rec else tuple(shp), rec, begin, recsize)

    def _i(self):
        v = struct.unpack_from(">i", self._b, self._p)[0]
        self._p += 4
        return v

    def _q(self):
        v = struct.unpack_from(">q", self._b, self._p)[0]
        self._p += 8
        return v

    def _s(self):
        n = self._i()
        s = self._b[self._p:self._p + n].decode()
        self._p += _pad4(n)
        return s

    def _atts(self):
        tag, n = self._i(), self._i()
        if tag not in (0, _ATT):
            raise CDFError("bad attribute list")
        out = {}
        for _ in range(n):
            nm = self._s()
            t, ne = self._i(), self._i()
            nb = ne * _SIZE[t]
            raw = self._b[self._p:self._p + nb]
            self._p += _pad4(nb)
            out[nm] = raw.decode(errors="replace") if t == NC_CHAR else np.frombuffer(raw, _NP[t]).astype(
                _NP[t][1:])
        return out


class Variable:
    def __init__(self, ds, name, dims, attrs, nctype, shape, record, begin, recsize):
        self.ds, self.name, self.dimensions, self.attrs = ds, name, dims, attrs
        self.nctype, self.record, self.begin, self.recsize = nctype, record, begin, recsize
        self.slab_shape = shape

    @property
    def shape(self):
        return ((self.ds.numrecs,) + self.slab_shape) if self.record else self.slab_shape

    def record_data(self, r):
        n = int(np.prod(self.slab_shape, dtype=np.int64))
        off = self.begin + r * self.recsize
        return np.frombuffer(self.ds._b, _NP[self.nctype], n, off).astype(_NP[self.nctype][1:]).reshape(
            self.slab_shape)

    def data(self):
        if self.record:
            return np.stack([self.record_data(r) for r in range(self.ds.numrecs)]) if self.ds.numrecs else \
                np.zeros((0,) + self.slab_shape, _NP[self.nctype][1:])
        n = int(np.prod(self.slab_shape, dtype=np.int64))
        return np.frombuffer(self.ds._b, _NP[self.nctype], n, self.begin).astype(_NP[self.nctype][1:]).reshape(
            self.slab_shape)


def read(path):
    return Dataset(path)


def file_size_ok(path):
    """True when the file is at least as long as its header promises."""
    ds = Dataset(path)
    end = 0
    for v in ds.variables.values():
        n = int(np.prod(v.slab_shape, dtype=np.int64)) * _SIZE[v.nctype]
        end = max(end, v.begin + (max(ds.numrecs - 1, 0) * v.recsize if v.record else 0) + n)
    return os.path.getsize(path) >= end
