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

_PT = np.dtype([("k", "<i4"), ("x", "<f8", (3,))])     # packed (key, Point3) record
_CX = np.dtype([("k", "<i4"), ("v", "<f8", (2,))])     # packed (key, cpx) record


def _write_parvec(path, rec, n):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", n))
        f.write(rec.tobytes())
        f.write(struct.pack("<iii", 2, 0, n))          # PtPrtn ownerinfo [0, N] (one rank)


def write_parvec_points(path, pos):
    """ParVec<int,Point3,PtPrtn> on one rank: map<int,Point3> (size, then key/value pairs)
    followed by the PtPrtn owner vector [0, N] (tt.cpp:66-82, parvec.hpp:530-550)."""
    pos = np.asarray(pos, dtype=np.float64)
    rec = np.empty(len(pos), dtype=_PT)
    rec["k"] = np.arange(len(pos))
    rec["x"] = pos
    _write_parvec(path, rec, len(pos))


def write_parvec_cpx(path, den):
    den = np.asarray(den, dtype=np.complex128)
    rec = np.empty(len(den), dtype=_CX)
    rec["k"] = np.arange(len(den))
    rec["v"] = den.view(np.float64).reshape(-1, 2)
    _write_parvec(path, rec, len(den))


def _read_parvec(path, dt):
    with open(path, "rb") as f:
        buf = f.read()
    n = struct.unpack_from("<i", buf, 0)[0]
    rec = np.frombuffer(buf, dtype=dt, count=n, offset=4)
    p = 4 + n * dt.itemsize
    m = struct.unpack_from("<i", buf, p)[0]
    own = list(struct.unpack_from(f"<{m}i", buf, p + 4))
    return rec, own


def read_parvec_cpx(path):
    rec, _ = _read_parvec(path, _CX)
    vals = rec["v"].copy().view(np.complex128).reshape(-1)
    return rec["k"].astype(np.int64), vals


def read_parvec_points(path):
    """Inverse of write_parvec_points: (keys, positions (n,3), ownerinfo)."""
    rec, own = _read_parvec(path, _PT)
    return rec["k"].astype(np.int64), rec["x"].copy(), own


def write_int_vec(path, v):
    """IntNumVec (serialize.hpp:360-382): m then m ints."""
    v = np.asarray(v, dtype=np.int32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(v)))
        f.write(v.tobytes())


def read_int_vec(path):
    with open(path, "rb") as f:
        r = _R(f.read())
    m = r.i32()
    return r._arr((m,), np.int32).astype(np.int64)


def write_int_tns(path, t):
    """IntNumTns (serialize.hpp:414-440): m, n, p then ints, first index fastest."""
    t = np.asarray(t, dtype=np.int32)
    with open(path, "wb") as f:
        for s in t.shape:
            f.write(struct.pack("<i", s))
        f.write(np.asfortranarray(t).tobytes(order="F"))


def read_int_tns(path):
    with open(path, "rb") as f:
        r = _R(f.read())
    m, n, p = r.i32(), r.i32(), r.i32()
    return r._arr((m, n, p), np.int32)
