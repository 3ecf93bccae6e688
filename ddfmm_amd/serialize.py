"""Reader/writer for the reference's little-endian binary formats.

The operator libraries (``<kind>3d_ld_<ACCU>.bin`` and ``<kind>3d_hd_<ACCU>_<NPQ>.bin``)
and the point/density/check files use the layout of include/serialize.hpp:24-711
(and matlab/serialize.m:18-124, which writes the same bytes):

* ``int``/``double`` raw 4/8 bytes, ``cpx`` = (re, im) doubles, ``Index3`` = 3 ints;
* ``DblNumMat`` = m, n (ints) then m*n doubles column-major (nummat.hpp:104-112);
* ``CpxNumMat``/``CpxNumTns`` likewise with complex entries;
* ``NumVec<T>`` = m then m items; ``NumTns<T>`` = m, n, p then items with the first
  index fastest (serialize.hpp:680-692);
* ``std::map<K,V>`` = size then (key, value) pairs in key order (serialize.hpp:214-229).

The low-frequency map is ``map<double, (uep, ucp, NumVec<CpxNumMat> uc2ue,
NumTns<CpxNumTns> ue2dc)>`` and the high-frequency map is
``map<double, map<Index3, (uep, ucp, NumVec<CpxNumMat> uc2ue)>>``
(mlib3d.hpp:27-55, mlib3d.cpp:361-405, matlab/aug3d_script.m:35-86).
"""
from __future__ import annotations

import io
import struct

import numpy as np


class _W:
    def __init__(self, f):
        self.f = f

    def i32(self, v):
        self.f.write(struct.pack("<i", int(v)))

    def f64(self, v):
        self.f.write(struct.pack("<d", float(v)))

    def dblmat(self, a):
        a = np.asarray(a, dtype=np.float64)
        self.i32(a.shape[0])
        self.i32(a.shape[1])
        self.f.write(np.asfortranarray(a).tobytes(order="F"))

    def cpxmat(self, a):
        a = np.asarray(a, dtype=np.complex128)
        if a.ndim == 1:
            a = a[:, None]
        self.i32(a.shape[0])
        self.i32(a.shape[1])
        self.f.write(np.asfortranarray(a).tobytes(order="F"))

    def cpxtns(self, a):
        if a is None:
            self.i32(0)
            self.i32(0)
            self.i32(0)
            return
        a = np.asarray(a, dtype=np.complex128)
        for s in a.shape:
            self.i32(s)
        self.f.write(np.asfortranarray(a).tobytes(order="F"))


class _R:
    def __init__(self, buf: bytes):
        self.b = memoryview(buf)
        self.p = 0

    def i32(self):
        v = struct.unpack_from("<i", self.b, self.p)[0]
        self.p += 4
        return v

    def f64(self):
        v = struct.unpack_from("<d", self.b, self.p)[0]
        self.p += 8
        return v

    def _arr(self, shape, dtype):
        n = int(np.prod(shape)) if len(shape) else 1
        nbytes = n * np.dtype(dtype).itemsize
        a = np.frombuffer(self.b[self.p:self.p + nbytes], dtype=dtype).copy()
        self.p += nbytes
        return a.reshape(shape, order="F")

    def dblmat(self):
        m, n = self.i32(), self.i32()
        return self._arr((m, n), np.float64)

    def cpxmat(self):
        m, n = self.i32(), self.i32()
        return self._arr((m, n), np.complex128)

    def cpxtns(self):
        m, n, p = self.i32(), self.i32(), self.i32()
        if m * n * p == 0:
            return None
        return self._arr((m, n, p), np.complex128)


def write_ld(path, ldmap):
    """Write ``{W: dict(uep, ucp, uc2ue=[V, invs, UH], ue2dc={(a,b,c): tensor})}``.

    ``ue2dc`` keys are offsets in [-3,3]^3 outside [-1,1]^3; stored at NumTns index
    offset+3 (mlib3d.cpp:162-171)."""
    with open(path, "wb") as f:
        w = _W(f)
        w.i32(len(ldmap))
        for W in sorted(ldmap):
            e = ldmap[W]
            w.f64(W)
            w.dblmat(e["uep"])
            w.dblmat(e["ucp"])
            w.i32(3)
            for m in e["uc2ue"]:
                w.cpxmat(m)
            w.i32(7)
            w.i32(7)
            w.i32(7)
            for c in range(7):
                for b in range(7):
                    for a in range(7):
                        w.cpxtns(e["ue2dc"].get((a - 3, b - 3, c - 3)))


def read_ld(path):
    with open(path, "rb") as f:
        r = _R(f.read())
    out = {}
    for _ in range(r.i32()):
        W = r.f64()
        uep = r.dblmat()
        ucp = r.dblmat()
        nm = r.i32()
        uc2ue = [r.cpxmat() for _ in range(nm)]
        m, n, p = r.i32(), r.i32(), r.i32()
        ue2dc = {}
        for c in range(p):
            for b in range(n):
                for a in range(m):
                    t = r.cpxtns()
                    if t is not None:
                        ue2dc[(a - 3, b - 3, c - 3)] = t
        out[W] = dict(uep=uep, ucp=ucp, uc2ue=uc2ue, ue2dc=ue2dc)
    return out


def write_hd(path, hdmap):
    """Write ``{W: {(a,b,c): dict(uep, ucp, uc2ue=[E1, E2, E3])}}``."""
    with open(path, "wb") as f:
        w = _W(f)
        w.i32(len(hdmap))
        for W in sorted(hdmap):
            w.f64(W)
            inner = hdmap[W]
            w.i32(len(inner))
            for key in sorted(inner):
                e = inner[key]
                for v in key:
                    w.i32(v)
                w.dblmat(e["uep"])
                w.dblmat(e["ucp"])
                w.i32(3)
                for m in e["uc2ue"]:
                    w.cpxmat(m)


def read_hd(path):
    with open(path, "rb") as f:
        r = _R(f.read())
    out = {}
    for _ in range(r.i32()):
        W = r.f64()
        inner = {}
        for _ in range(r.i32()):
            key = (r.i32(), r.i32(), r.i32())
            uep = r.dblmat()
            ucp = r.dblmat()
            nm = r.i32()
            inner[key] = dict(uep=uep, ucp=ucp, uc2ue=[r.cpxmat() for _ in range(nm)])
        out[W] = inner
    return out


# ---- point / density / check-key files (tt.cpp:66-98, 169-177; new_data.m:158-242) ----

def write_parvec_points(path, pos):
    """ParVec<int,Point3,PtPrtn> on one rank: map<int,Point3> then PtPrtn ownerinfo.

    Mirrors the per-rank file read by tt.cpp:66-82 (deserialize of the ParVec's local
    map followed by the partition vector ``[0, N]``)."""
    pos = np.asarray(pos, dtype=np.float64)
    buf = io.BytesIO()
    w = _W(buf)
    w.i32(len(pos))
    for k, p in enumerate(pos):
        w.i32(k)
        for d in range(3):
            w.f64(p[d])
    w.i32(2)
    w.i32(0)
    w.i32(len(pos))
    with open(path, "wb") as f:
        f.write(buf.getvalue())


def write_parvec_cpx(path, den):
    den = np.asarray(den, dtype=np.complex128)
    buf = io.BytesIO()
    w = _W(buf)
    w.i32(len(den))
    for k, v in enumerate(den):
        w.i32(k)
        w.f64(v.real)
        w.f64(v.imag)
    w.i32(2)
    w.i32(0)
    w.i32(len(den))
    with open(path, "wb") as f:
        f.write(buf.getvalue())


def read_parvec_cpx(path):
    with open(path, "rb") as f:
        r = _R(f.read())
    n = r.i32()
    keys = np.empty(n, dtype=np.int64)
    vals = np.empty(n, dtype=np.complex128)
    for k in range(n):
        keys[k] = r.i32()
        vals[k] = complex(r.f64(), r.f64())
    return keys, vals
