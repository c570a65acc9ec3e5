"""A streaming writer of netCDF classic files (64-bit offset variant), the
format of the reference's output (netCDF-2 calls nccre/ncddef/ncvdef/ncvpt,
write_snapshota.f, put_block.f, put_block1.f).

The header is written when the file is created, from the full list of
dimensions, attributes and variables (the reference's define mode ends with
ncendf before any data).  Fixed-size variables are written with put(); each
append() writes one record (one slice of every record variable along the
unlimited dimension) and updates the record count, so a run of any length
streams to disk without holding its records in memory.

Layout (the published netCDF classic format): magic 'CDF' + version byte 2,
record count, dimension list, global attributes, variable list (name,
dimension ids, attributes, type, per-variable size and 8-byte begin offset),
then the fixed variables' data, then the records.  Big-endian throughout;
names, attribute values and variable sizes padded to 4 bytes."""
import struct

import numpy as np

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 10, 11, 12
NC_TYPES = {np.dtype("int8"): 1, np.dtype("S1"): 2, np.dtype("int16"): 3, np.dtype("int32"): 4,
            np.dtype("float32"): 5, np.dtype("float64"): 6}


def _pad4(n):
    return (n + 3) & ~3


def _name(s):
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attr_value(v):
    """(nc_type, nelems, padded big-endian bytes) of an attribute value: str
    -> NC_CHAR, numpy scalar/array -> its own type, Python float -> double,
    Python int -> int."""
    if isinstance(v, str):
        b = v.encode()
        return 2, len(b), b + b"\0" * (_pad4(len(b)) - len(b))
    a = np.atleast_1d(np.asarray(v))
    if a.dtype.kind == "i" and not isinstance(v, (np.ndarray, np.generic)):
        a = a.astype(np.int32)
    t = NC_TYPES[a.dtype]
    b = a.astype(a.dtype.newbyteorder(">")).tobytes()
    return t, a.size, b + b"\0" * (_pad4(len(b)) - len(b))


def _attrs(d):
    if not d:
        return struct.pack(">ii", 0, 0)
    out = struct.pack(">ii", NC_ATTRIBUTE, len(d))
    for k, v in d.items():
        t, n, b = _attr_value(v)
        out += _name(k) + struct.pack(">ii", t, n) + b
    return out


class NcWriter:
    """dims: [(name, length or None for the unlimited one)];
    variables: [(name, dtype, (dim names, slowest first), {attributes})]."""

    def __init__(self, path, dims, variables, global_attrs=None):
        self.dims = [(n, l) for n, l in dims]
        if sum(l is None for _, l in self.dims) > 1:
            raise ValueError("at most one unlimited dimension")
        self.dimlen = {n: (0 if l is None else int(l)) for n, l in self.dims}
        self.dimid = {n: q for q, (n, _) in enumerate(self.dims)}
        self.recdim = next((n for n, l in self.dims if l is None), None)
        self.vars = {}
        order = []
        for name, dtype, vdims, atts in variables:
            dt = np.dtype(dtype)
            if dt not in NC_TYPES:
                raise ValueError("unsupported type %s" % dt)
            rec = bool(vdims) and vdims[0] == self.recdim
            if any(d == self.recdim for d in vdims[1:]):
                raise ValueError("the unlimited dimension must come first")
            shape = tuple(self.dimlen[d] for d in (vdims[1:] if rec else vdims))
            nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
            self.vars[name] = {"dtype": dt, "dims": tuple(vdims), "atts": dict(atts or {}), "rec": rec,
                               "shape": shape, "nbytes": nbytes, "vsize": _pad4(nbytes)}
            order.append(name)
        self.order = order
        self.gatts = dict(global_attrs or {})
        self.numrecs = 0
        # begin offsets need the header size, which depends on them only
        # through their fixed 8-byte width
        hdr = len(self._header())
        off = hdr
        for n in order:
            v = self.vars[n]
            if not v["rec"]:
                v["begin"] = off
                off += v["vsize"]
        self.recstart = off
        recvars = [n for n in order if self.vars[n]["rec"]]
        # a lone record variable is not padded (the format's special case)
        if len(recvars) == 1:
            self.vars[recvars[0]]["vsize"] = self.vars[recvars[0]]["nbytes"]
        self.recsize = 0
        for n in recvars:
            self.vars[n]["begin"] = self.recstart + self.recsize
            self.recsize += self.vars[n]["vsize"]
        self.fh = open(path, "w+b")
        h = self._header()
        assert len(h) == hdr
        self.fh.write(h)
        # fixed variables start as zeros until put()
        self.fh.truncate(self.recstart)

    def _header(self):
        out = b"CDF\x02" + struct.pack(">i", self.numrecs)
        if self.dims:
            out += struct.pack(">ii", NC_DIMENSION, len(self.dims))
            for n, _ in self.dims:
                out += _name(n) + struct.pack(">i", self.dimlen[n])
        else:
            out += struct.pack(">ii", 0, 0)
        out += _attrs(self.gatts)
        if self.order:
            out += struct.pack(">ii", NC_VARIABLE, len(self.order))
            for n in self.order:
                v = self.vars[n]
                out += _name(n) + struct.pack(">i", len(v["dims"]))
                out += b"".join(struct.pack(">i", self.dimid[d]) for d in v["dims"])
                out += _attrs(v["atts"])
                # a slice over 2^32 - 4 bytes records vsize as 2^32 - 1
                vs = v["vsize"] if v["vsize"] < 2**32 - 4 else 2**32 - 1
                out += struct.pack(">iIq", NC_TYPES[v["dtype"]], vs, v.get("begin", 0))
        else:
            out += struct.pack(">ii", 0, 0)
        return out

    def _bytes(self, name, data):
        v = self.vars[name]
        a = np.asarray(data, dtype=v["dtype"])
        if a.size != int(np.prod(v["shape"], dtype=np.int64)):
            raise ValueError("%s: %d values for shape %s" % (name, a.size, v["shape"]))
        # data are shaped like the reference's Fortran arrays (x first): read
        # x fastest, which is the netCDF order with the dimensions reversed
        b = np.ravel(a, order="F").astype(v["dtype"].newbyteorder(">")).tobytes()
        return b + b"\0" * (v["vsize"] - len(b))

    def put(self, name, data):
        """Write a fixed-size variable; data shaped like the reference's
        Fortran array (x first), or flat with x fastest."""
        v = self.vars[name]
        if v["rec"]:
            raise ValueError("%s is a record variable: use append" % name)
        self.fh.seek(v["begin"])
        self.fh.write(self._bytes(name, data))

    def append(self, record):
        """Write record number numrecs: record maps every record variable to
        its slice (x fastest)."""
        recvars = [n for n in self.order if self.vars[n]["rec"]]
        missing = [n for n in recvars if n not in record]
        if missing:
            raise ValueError("record lacks %s" % missing)
        base = self.recstart + self.numrecs * self.recsize
        for n in recvars:
            self.fh.seek(base + (self.vars[n]["begin"] - self.recstart))
            self.fh.write(self._bytes(n, record[n]))
        self.numrecs += 1
        self.fh.seek(4)
        self.fh.write(struct.pack(">i", self.numrecs))

    def close(self):
        if self.fh:
            self.fh.flush()
            self.fh.close()
            self.fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
