"""Reader for NetCDF files in the classic and 64-bit-offset formats.

These are the formats WRF writes with io_form 2 when its netCDF library does
not compress (the case's wrfinput / wrfbdy / wrfout files).  The reference
opens them through xarray and netCDF4 (solver.py:40-45); this reader needs
neither: it parses the header and exposes every variable as a numpy view of
a read-only memory map of the file, so a boundary record is read only when
it is used.  netCDF-4 (HDF5) files are refused with an error that says so.

Layout (the NetCDF classic format specification): the magic "CDF" and a
version byte (1: 32-bit offsets, 2: 64-bit offsets), the record count, then
the dimension, global attribute and variable lists, each tagged
(NC_DIMENSION 10, NC_ATTRIBUTE 12, NC_VARIABLE 11; an absent list is two
zero words).  Names are counted strings padded to 4 bytes; all numbers are
big-endian.  A variable whose first dimension is the record dimension is
stored record by record, one slab of each record variable per record (the
record size is the sum of their padded slab sizes, or the one variable's
unpadded slab size when it is alone).
"""
import mmap

import numpy as np

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12
_TYPES = {1: np.dtype("i1"), 2: np.dtype("S1"), 3: np.dtype(">i2"), 4: np.dtype(">i4"), 5: np.dtype(">f4"),
          6: np.dtype(">f8")}


class Variable:
    """One variable: name, dimension names, shape, dtype, attributes, and its
    data as a numpy array (a view of the file; `[...]` copies a selection
    into native byte order)."""

    def __init__(self, name, dims, shape, dtype, attrs, data):
        self.name, self.dims, self.shape, self.dtype, self.attrs = name, dims, shape, dtype, attrs
        self.data = data

    def __getitem__(self, key):
        a = self.data[key]
        return np.ascontiguousarray(a).astype(a.dtype.newbyteorder("="), copy=False)

    def __repr__(self):
        return "<Variable %s%s %s>" % (self.name, self.dims, self.dtype)


class Dataset:
    """A NetCDF classic / 64-bit-offset file, read-only."""

    def __init__(self, path):
        self.path = path
        self._f = open(path, "rb")
        try:
            self._mm = mmap.mmap(self._f.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError:                       # an empty file
            raise ValueError("%s: not a NetCDF file (empty)" % path)
        self._pos = 0
        magic = self._bytes(4)
        if magic[:3] != b"CDF" or magic[3] not in (1, 2):
            if magic[:4] == b"\x89HDF":
                raise ValueError("%s: a netCDF-4 (HDF5) file; convert it to the classic or 64-bit-offset "
                                 "format (nccopy -k classic / -k 64-bit-offset)" % path)
            raise ValueError("%s: not a NetCDF classic or 64-bit-offset file" % path)
        self.version = magic[3]
        self.numrecs = self._u32()
        self.dimensions = {}
        self._dimlist = []
        self._unlimited = None
        for name, n in self._list(_NC_DIMENSION, self._dim):
            if n == 0:
                self._unlimited = name
            self.dimensions[name] = n
            self._dimlist.append(name)
        self.attrs = dict(self._list(_NC_ATTRIBUTE, self._attr))
        raw = self._list(_NC_VARIABLE, self._var)
        recs = [v for v in raw if v["record"]]
        if len(recs) == 1:
            recsize = recs[0]["vsize"]
        else:
            recsize = sum(v["vsize"] for v in recs)
        self.variables = {}
        for v in raw:
            dt = v["dtype"]
            inner = tuple(self.dimensions[d] for d in v["dims"][1:]) if v["record"] else \
                tuple(self.dimensions[d] for d in v["dims"])
            inner_strides = tuple(int(np.prod(inner[k + 1:], dtype=np.int64)) * dt.itemsize for k in range(len(inner)))
            if v["record"]:
                shape = (self.numrecs,) + inner
                strides = (recsize,) + inner_strides
            else:
                shape, strides = inner, inner_strides
            need = v["begin"] + (0 if not shape or 0 in shape else
                                 sum((n - 1) * st for n, st in zip(shape, strides)) + dt.itemsize)
            if need > len(self._mm):
                raise ValueError("%s: variable %s reaches past the end of the file" % (path, v["name"]))
            data = np.ndarray(shape, dtype=dt, buffer=self._mm, offset=v["begin"], strides=strides)
            self.variables[v["name"]] = Variable(v["name"], v["dims"], shape, dt, v["attrs"], data)

    # -- header parsing
    def _bytes(self, n):
        b = self._mm[self._pos:self._pos + n]
        if len(b) != n:
            raise ValueError("%s: truncated header" % self.path)
        self._pos += n
        return b

    def _u32(self):
        return int.from_bytes(self._bytes(4), "big")

    def _u64(self):
        return int.from_bytes(self._bytes(8), "big")

    def _name(self):
        n = self._u32()
        s = self._bytes(n).decode("utf-8")
        self._bytes((4 - n % 4) % 4)
        return s

    def _list(self, tag, item):
        t, n = self._u32(), self._u32()
        if t == 0 and n == 0:
            return []
        if t != tag:
            raise ValueError("%s: malformed header (tag %d, expected %d)" % (self.path, t, tag))
        return [item() for _ in range(n)]

    def _dim(self):
        return self._name(), self._u32()

    def _attr(self):
        name = self._name()
        nc_type, n = self._u32(), self._u32()
        dt = _TYPES.get(nc_type)
        if dt is None:
            raise ValueError("%s: attribute %s has unknown type %d" % (self.path, name, nc_type))
        nbytes = n * dt.itemsize
        raw = self._bytes(nbytes)
        self._bytes((4 - nbytes % 4) % 4)
        if nc_type == 2:
            return name, raw.rstrip(b"\0").decode("utf-8", "replace")
        val = np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))
        return name, (val[0].item() if n == 1 else val)

    def _var(self):
        name = self._name()
        ndims = self._u32()
        dims = tuple(self._dimlist[self._u32()] for _ in range(ndims))
        attrs = dict(self._list(_NC_ATTRIBUTE, self._attr))
        nc_type = self._u32()
        dt = _TYPES.get(nc_type)
        if dt is None:
            raise ValueError("%s: variable %s has unknown type %d" % (self.path, name, nc_type))
        vsize = self._u32()
        begin = self._u64() if self.version == 2 else self._u32()
        record = bool(dims) and dims[0] == self._unlimited
        return dict(name=name, dims=dims, attrs=attrs, dtype=dt, vsize=vsize, begin=begin, record=record)

    # -- access
    def __getitem__(self, name):
        return self.variables[name]

    def __contains__(self, name):
        return name in self.variables

    def close(self):
        self.variables = {}
        try:
            self._mm.close()
        except BufferError:           # arrays viewing the map are still alive; it closes with them
            pass
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
