"""ORACLE (test infrastructure only) — a CPU FP64 restatement of the reference's
directional FMM (``Wave3d::setup`` + ``Wave3d::eval``) in numpy.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use
this module, and only as the checker.  It follows, function by function:

* ``Kernel3d::kernel``           kernel3d.cpp:26-107 (r^2 < 1e-16 => r := 1, k = 2*pi)
* ``nml2dir/ParentDir/ChildDir``  wave3d.cpp:43-119; ``Dir2Width`` 122-128; ``BoxCenter`` 131-142
* ``setup_tree``                 wave3d_setup.cpp:100-299 (split rule 156-157, child bit 167,
                                 direction sets 254-296)
* ``setup_tree_callowlist``      wave3d_setup.cpp:302-390 (U/V/W/X lists)
* ``setup_tree_calhghlist``      wave3d_setup.cpp:393-443 (far lists keyed target - source)
* ``setup_tree_find/adjacent``   wave3d_setup.cpp:446-477
* ``Mlib3d::*Fetch``             mlib3d.cpp:106-357 (symmetry sort/shuffle, child kernels)
* ``eval``                       wave3d_eval.cpp:299-415, ``ConstructMaps`` 252-295,
  ``EvalUpwardLow`` 418-490, ``EvalDownwardLow`` 493-583, U/V/X/W lists 585-718,
  HF recursion 721-751, ``EvalUpwardHigh`` 773-854, ``HighFrequencyM2L`` 876-928,
  ``HighFrequencyL2L`` 931-993, ``EvalDownwardHigh`` 996-1040
* ``check``                      wave3d_check.cpp:22-100 (direct sum at check keys)

Deviation kept on purpose: the HF M2L matrix G(dcp + c_t, uep + c_s) is evaluated once
per (W, c_t - c_s) from the centre difference instead of per pair from absolute
positions; this changes only rounding (~1e-15 relative).
"""
from __future__ import annotations

import ctypes
import math
import os
from collections import deque

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
TPK = 2.0 * math.pi
MINDIF2 = 1e-8 * 1e-8
PTS, TERMINAL = 1, 2


def kernel(trg: np.ndarray, src: np.ndarray, ktype: int) -> np.ndarray:
    """G(trg_i, src_j) for (m,3) and (n,3) point rows, kernel3d.cpp:26-107."""
    d = trg[:, None, :] - src[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    r2 = np.where(r2 < MINDIF2, 1.0, r2)
    r = np.sqrt(r2)
    kr = r * TPK
    if ktype == 0:
        ir = 1.0 / r
        return np.cos(kr) * ir + 1j * (np.sin(kr) * ir)
    if ktype == 1:
        return np.cos(kr) + 1j * np.sin(kr)
    raise ValueError(f"Unknown kernel type {ktype}")


# ---------------------------------------------------------------- C direct sum
_lib = None


def _direct_lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libdirect_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libdirect_ref.so missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.direct_sum.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int]
        _lib = lib
    return _lib


def direct_sum(trg: np.ndarray, src: np.ndarray, den: np.ndarray, ktype: int = 0,
               threads: int = 0) -> np.ndarray:
    """FP64 direct sum u(t) = sum_j G(t, s_j) f_j with the same coincidence rule
    (wave3d_check.cpp:66-76); den is (n,) or (nrhs, n); uses the C helper."""
    lib = _direct_lib()
    trg = np.ascontiguousarray(trg, dtype=np.float64)
    src = np.ascontiguousarray(src, dtype=np.float64)
    den2 = np.ascontiguousarray(np.atleast_2d(den), dtype=np.complex128)
    nrhs = den2.shape[0]
    out = np.zeros((nrhs, len(trg)), dtype=np.complex128)
    lib.direct_sum(len(trg), trg.ctypes.data, src.ctypes.data, len(src),
                   den2.ctypes.data, nrhs, out.ctypes.data, None, int(ktype) | (int(threads) << 8))
    return out[0] if np.ndim(den) == 1 else out


def relerr(approx: np.ndarray, exact: np.ndarray) -> float:
    return float(np.linalg.norm(approx - exact) / np.linalg.norm(exact))


def _translate_fp64(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    return A @ x


# Every translation matvec of eval (LF M2M/L2L, HF U2U/M2L/D2D; the reference's zgemv
# calls in wave3d_eval.cpp:461, 578, 812, 831, 925, 967, 989) goes through
# this hook, so that tests can substitute an emulated reduced-precision product; the
# check-to-equivalent solves, the kernel applications and the FFT stay FP64.
translate = _translate_fp64


# ---------------------------------------------------------------- direction algebra
def nml2dir(n, W: float, NPQ: int):
    """wave3d.cpp:43-71."""
    C = NPQ * int(round(W))
    B = C // 2
    midx = 0
    mval = abs(n[0])
    for d in (1, 2):
        if mval < abs(n[d]):
            midx = d
            mval = abs(n[d])
    val = [math.atan(n[d] / mval) for d in range(3)]
    Ang = (math.pi / 2) / C
    res = []
    for d in range(3):
        tmp = int(math.floor(val[d] / Ang))
        tmp = min(max(tmp, -B), B - 1)
        res.append(2 * tmp + 1)
    res[midx] = C * int(_cround(val[midx]))
    return tuple(res)


def _cround(x: float) -> float:
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _cdiv(a: int, b: int) -> int:
    """C++ integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def parent_dir(d):
    """wave3d.cpp:74-95 (identical to Mlib3d::predir, mlib3d.cpp:283-303)."""
    C = max(abs(v) for v in d)
    B = C // 2
    midx = -1
    for k in range(3):
        if abs(d[k]) == C:
            midx = k
    res = []
    for k in range(3):
        r = _cdiv(d[k] + C - 1, 2)
        r = 2 * _cdiv(r, 2) + 1 - B
        res.append(r)
    res[midx] = _cdiv(d[midx], 2)
    return tuple(res)


def child_dirs(d):
    """wave3d.cpp:98-119."""
    C = max(abs(v) for v in d)
    oidx = [k for k in range(3) if abs(d[k]) != C]
    out = []
    for a in range(2):
        for b in range(2):
            t = [2 * v for v in d]
            t[oidx[0]] += 2 * a - 1
            t[oidx[1]] += 2 * b - 1
            out.append(tuple(t))
    return out


def sort_dir(val):
    """Mlib3d::HighFetchIndex3Sort (mlib3d.cpp:328-357): returns (srt, sgn, prm)."""
    sgn = [1 if v >= 0 else -1 for v in val]
    v = [abs(x) for x in val]
    upp = max(v) + 1
    srt, prm = [0, 0, 0], [0, 0, 0]
    for d in range(3):
        minidx, minval = d, v[d]
        for e in range(3):
            if v[e] < minval:
                minval, minidx = v[e], e
        srt[d], prm[d] = minval, minidx
        v[minidx] = upp
    return tuple(srt), tuple(sgn), tuple(prm)


def shuffle(prm, sgn, pts3xn: np.ndarray) -> np.ndarray:
    """Mlib3d::HighFetchShuffle (mlib3d.cpp:308-325); returns (n,3) rows."""
    res = np.empty_like(pts3xn)
    for d in range(3):
        res[prm[d], :] = pts3xn[d, :]
    for d in range(3):
        res[d, :] *= sgn[d]
    return res.T.copy()


def shifted_point(ind: int, W: float):
    a, b, c = (ind & 4) >> 2, (ind & 2) >> 1, ind & 1
    return np.array([(a - 0.5) * W / 2, (b - 0.5) * W / 2, (c - 0.5) * W / 2])


def child_key(key, ind):
    l, i, j, k = key
    return (l + 1, 2 * i + ((ind & 4) >> 2), 2 * j + ((ind & 2) >> 1), 2 * k + (ind & 1))


def parent_key(key):
    l, i, j, k = key
    return (l - 1, i // 2, j // 2, k // 2)


class Box:
    __slots__ = ("tag", "ptidx", "U", "V", "W", "X", "E", "fnde", "incdir", "outdir",
                 "fftnum")

    def __init__(self, ptidx):
        self.tag = 0
        self.ptidx = ptidx
        self.U, self.V, self.W, self.X, self.E = [], [], [], [], []
        self.fnde = {}
        self.incdir, self.outdir = set(), set()
        self.fftnum = 0


# ---------------------------------------------------------------- Wave3d
class Wave3dOracle:
    def __init__(self, pos, K, ld, hd, ktype=0, ACCU=2, NPQ=4, ctr=(0.0, 0.0, 0.0),
                 ptsmax=100, maxlevel=12, numC=1):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.K = float(K)
        self.ld, self.hd = ld, hd
        self.ktype, self.ACCU, self.NPQ = ktype, ACCU, NPQ
        self.ctr = np.asarray(ctr, dtype=np.float64)
        self.ptsmax, self.maxlevel, self.numC = ptsmax, maxlevel, numC
        self.P = {1: 4, 2: 6, 3: 8}[ACCU]

    # -- geometry
    def unitlevel(self):
        return int(round(math.log(self.K) / math.log(2)))

    def cell_level(self):
        return int(round(math.log(self.numC) / math.log(2)))

    def width(self, key):
        return self.K / (1 << key[0])

    def center(self, key):
        l, i, j, k = key
        tmp = 1 << l
        return np.array([self.ctr[d] - self.K / 2 + (idx + 0.5) / tmp * self.K
                         for d, idx in enumerate((i, j, k))])

    def has_pts(self, b):
        return b.tag & PTS

    def terminal(self, b):
        return b.tag & TERMINAL

    # -- setup (wave3d_setup.cpp:100-299)
    def setup(self):
        K, numC, lvlC = self.K, self.numC, self.cell_level()
        idx = np.floor(numC * ((self.pos - self.ctr + K / 2) / K)).astype(np.int64)
        if np.any(idx < 0) or np.any(idx >= numC):
            raise ValueError("point outside the computational domain")
        cells = {}
        for p in range(len(self.pos)):
            cells.setdefault(tuple(idx[p]), []).append(p)
        q = deque()
        for a in range(numC):
            for b in range(numC):
                for c in range(numC):
                    q.append(((lvlC, a, b, c), Box(np.array(cells.get((a, b, c), []), dtype=np.int64))))
        boxes = {}
        ul = self.unitlevel()
        while q:
            key, box = q.popleft()
            npts = len(box.ptidx)
            if npts > 0:
                box.tag |= PTS
            action = (key[0] <= ul and npts > 0) or (npts > self.ptsmax and key[0] < self.maxlevel - 1)
            if action:
                cc = self.center(key)
                p = self.pos[box.ptidx]
                bits = (p >= cc).astype(np.int64)
                code = bits[:, 0] * 4 + bits[:, 1] * 2 + bits[:, 2]
                for ind in range(8):
                    q.append((child_key(key, ind), Box(box.ptidx[code == ind])))
                box.ptidx = np.zeros(0, dtype=np.int64)
            else:
                box.tag |= TERMINAL
            boxes[key] = box
        self.boxes = dict(sorted(boxes.items()))
        eps = 1e-12
        for key, box in self.boxes.items():
            if self.has_pts(box):
                if self.width(key) < 1 - eps:
                    self._callowlist(key, box)
                else:
                    self._calhghlist(key, box)
        for key, box in self.boxes.items():
            if self.has_pts(box):
                for nk in box.V:
                    self.boxes[nk].fftnum += 1
        for key, box in self.boxes.items():
            W = self.width(key)
            if W > 1 - eps and self.has_pts(box):
                if key[0] != lvlC:
                    par = self.boxes[parent_key(key)]
                    for d in par.outdir:
                        box.outdir.add(parent_dir(d))
                    for d in par.incdir:
                        box.incdir.add(parent_dir(d))
                cc = self.center(key)
                for lst in box.fnde.values():
                    for ok in lst:
                        oc = self.center(ok)
                        t = oc - cc
                        t = t / math.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
                        box.outdir.add(nml2dir(t, W, self.NPQ))
                        t = cc - oc
                        t = t / math.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
                        box.incdir.add(nml2dir(t, W, self.NPQ))
        return self

    def _find(self, key):
        lvlC = self.cell_level()
        k = key
        while k[0] != lvlC:
            if k in self.boxes:
                return True, k
            k = parent_key(k)
        return (k in self.boxes), k

    @staticmethod
    def adjacent(me, yo):
        md = max(me[0], yo[0])
        sm, sy = 1 << (md - me[0]), 1 << (md - yo[0])
        rad = sm + sy
        dif = [abs((2 * me[d + 1] + 1) * sm - (2 * yo[d + 1] + 1) * sy) for d in range(3)]
        return all(x <= rad for x in dif) and max(dif) == rad

    def _callowlist(self, key, box):
        lvlC = self.cell_level()
        if key[0] == lvlC:
            raise ValueError("cell-level boxes must have width >= 1")
        U, V, Wl, X = set(), set(), set(), set()
        par = parent_key(key)
        L = 1 << key[0]
        rng = [range(max(2 * par[d + 1] - 2, 0), min(2 * par[d + 1] + 4, L)) for d in range(3)]
        term = self.terminal(box)
        for i in rng[0]:
            for j in rng[1]:
                for k in rng[2]:
                    if (i, j, k) == key[1:]:
                        continue
                    found, res = self._find((key[0], i, j, k))
                    if not found:
                        continue
                    rb = self.boxes[res]
                    adj = self.adjacent(res, key)
                    if res[0] < key[0] and self.has_pts(rb):
                        if not adj:
                            X.add(res)
                        elif term:
                            U.add(res)
                    if res[0] == key[0]:
                        if not adj:
                            assert max(abs(res[d] - key[d]) for d in (1, 2, 3)) <= 3
                            if self.has_pts(rb):
                                V.add(res)
                        elif term:
                            rest = deque([res])
                            while rest:
                                fk = rest.popleft()
                                fb = self.boxes[fk]
                                a2 = self.adjacent(fk, key)
                                if not a2 and self.has_pts(fb):
                                    Wl.add(fk)
                                if a2 and self.terminal(fb) and self.has_pts(fb):
                                    U.add(fk)
                                if a2 and not self.terminal(fb):
                                    for ind in range(8):
                                        rest.append(child_key(fk, ind))
        if term and self.has_pts(box):
            U.add(key)
        box.U, box.V, box.W, box.X = sorted(U), sorted(V), sorted(Wl), sorted(X)

    def _calhghlist(self, key, box):
        cc = self.center(key)
        W = self.width(key)
        thr = W * W + W - 1e-12
        lvlC = self.cell_level()
        if key[0] == lvlC:
            cands = [k for k in self.boxes if k[0] == lvlC]
        else:
            cands = [child_key(tk, ind) for tk in self.boxes[parent_key(key)].E for ind in range(8)]
        for ok in cands:
            if self.has_pts(self.boxes[ok]):
                diff = cc - self.center(ok)
                nrm = math.sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2])
                if nrm >= thr:
                    box.fnde.setdefault(nml2dir(diff / nrm, W, self.NPQ), []).append(ok)
                else:
                    box.E.append(ok)

    # -- operator fetch (mlib3d.cpp:106-280)
    def _lf(self, W):
        e = self.ld[W]
        return e["uep"].T, e["ucp"].T, e["uc2ue"], e["ue2dc"]

    def _hf_pts(self, W, d):
        srt, sgn, prm = sort_dir(d)
        e = self.hd[float(W)][srt]
        return shuffle(prm, sgn, e["uep"]), shuffle(prm, sgn, e["ucp"]), e["uc2ue"]

    def _hf_child_uep(self, W, d):
        if W == 1:
            return self._lf(W / 2)[0]
        return self._hf_pts(W / 2, parent_dir(d))[0]

    # -- eval (wave3d_eval.cpp:299-415)
    def eval(self, den: np.ndarray) -> np.ndarray:
        den = np.asarray(den, dtype=np.complex128)
        kt = self.ktype
        eps = 1e-12
        boxes = {k: b for k, b in self.boxes.items() if self.has_pts(b)}
        self.live = boxes
        extpos = {k: self.pos[b.ptidx] for k, b in boxes.items() if self.terminal(b)}
        extden = {k: den[b.ptidx] for k, b in boxes.items() if self.terminal(b)}
        upeq, dnchk, extval = {}, {}, {}
        ldmap, hdmap = {}, {}
        for k, b in boxes.items():
            W = self.width(k)
            if W < 1 - eps:
                ldmap.setdefault(W, []).append(k)
            else:
                for d in sorted(b.outdir):
                    hdmap.setdefault(d, ([], []))[0].append(k)
                for d in sorted(b.incdir):
                    hdmap.setdefault(d, ([], []))[1].append(k)
        hdmap = dict(sorted(hdmap.items()))
        # LF upward, finest first (wave3d_eval.cpp:50-68, 418-490)
        for W in sorted(ldmap):
            uep, ucp, (V, invs, UH), _ = self._lf(W)
            uepc = self._lf(W / 2)[0] if (W / 2) in self.ld else None
            ue2uc = None
            for k in ldmap[W]:
                b = boxes[k]
                c = self.center(k)
                if self.terminal(b):
                    chk = kernel(ucp + c, extpos[k], kt) @ extden[k]
                else:
                    if ue2uc is None:
                        ue2uc = [kernel(ucp, uepc + shifted_point(i, W), kt) for i in range(8)]
                    chk = np.zeros(len(ucp), dtype=np.complex128)
                    for ind in range(8):
                        ck = child_key(k, ind)
                        if ck in boxes:
                            chk = chk + translate(ue2uc[ind], upeq[ck])
                upeq[k] = V @ ((UH @ chk) * invs[:, 0])
        # HF pass (wave3d_eval.cpp:188-235)
        dirup, dirdn = {}, {}
        base = [d for d in hdmap if max(abs(v) for v in d) // self.NPQ == 1]

        def up_rec(W, d):
            if d not in hdmap:
                return
            self._up_high(W, d, hdmap[d], boxes, upeq, dirup)
            for cd in child_dirs(d):
                up_rec(2 * W, cd)

        def dn_rec(W, d):
            if d not in hdmap:
                return
            for cd in child_dirs(d):
                dn_rec(2 * W, cd)
            self._down_high(W, d, hdmap[d], boxes, dirup, dirdn, dnchk)

        for d in base:
            up_rec(1, d)
        for d in base:
            dn_rec(1, d)
        # LF downward, coarsest first (wave3d_eval.cpp:92-104, 493-583)
        fftcache = {}
        dneq_all = {}
        for W in sorted(ldmap, reverse=True):
            uep, ucp, (V, invs, UH), ue2dc = self._lf(W)
            dep, dcp = ucp, uep
            dcpc = self._lf(W / 2)[0] if (W / 2) in self.ld else None
            de2dc = None
            P = self.P
            step = W / (P - 1)
            gi_e = np.rint((uep + W / 2) / step).astype(np.int64) + P
            for k in ldmap[W]:
                b = boxes[k]
                c = self.center(k)
                chk = dnchk.get(k, np.zeros(len(dcp), dtype=np.complex128))
                term = self.terminal(b)
                ev = np.zeros(len(b.ptidx), dtype=np.complex128)
                tp = extpos.get(k)
                for nk in b.U:
                    ev = ev + kernel(tp, extpos[nk], kt) @ extden[nk]
                # V list with FFT-diagonal translation
                acc = np.zeros((2 * P,) * 3, dtype=np.complex128)
                for nk in b.V:
                    off = tuple(int(_cround(x)) for x in (c - self.center(nk)) / W)
                    if nk not in fftcache:
                        g = np.zeros((2 * P,) * 3, dtype=np.complex128)
                        g[gi_e[:, 0], gi_e[:, 1], gi_e[:, 2]] = upeq[nk]
                        fftcache[nk] = np.fft.fftn(g)
                    acc += fftcache[nk] * ue2dc[off]
                if b.V:
                    vals = np.fft.ifftn(acc)
                    chk = chk + vals[gi_e[:, 0], gi_e[:, 1], gi_e[:, 2]]
                for nk in b.W:
                    nb = boxes[nk]
                    if self.terminal(nb) and len(nb.ptidx) < len(uep):
                        ev = ev + kernel(tp, extpos[nk], kt) @ extden[nk]
                    else:
                        coef = self.width(nk) / W
                        ev = ev + kernel(tp, coef * uep + self.center(nk), kt) @ upeq[nk]
                for nk in b.X:
                    if term and len(b.ptidx) < len(dcp):
                        ev = ev + kernel(tp, extpos[nk], kt) @ extden[nk]
                    else:
                        chk = chk + kernel(dcp + c, extpos[nk], kt) @ extden[nk]
                dneq = UH.T @ ((V.T @ chk) * invs[:, 0])
                dneq_all[k] = dneq
                if term:
                    ev = ev + kernel(tp, dep + c, kt) @ dneq
                    extval[k] = ev
                else:
                    if de2dc is None:
                        de2dc = [kernel(dcpc + shifted_point(i, W), dep, kt) for i in range(8)]
                    for ind in range(8):
                        ck = child_key(k, ind)
                        if ck in boxes:
                            dnchk[ck] = dnchk.get(ck, 0) + translate(de2dc[ind], dneq)
        val = np.zeros(len(self.pos), dtype=np.complex128)
        for k, ev in extval.items():
            val[boxes[k].ptidx] = ev
        # intermediate quantities, kept for stage-by-stage comparisons
        self.upeq, self.dirup, self.dirdn, self.dneq, self.hdmap = upeq, dirup, dirdn, dneq_all, hdmap
        return val

    def _up_high(self, W, d, lists, boxes, upeq, dirup):
        uep, ucp, (E1, E2, E3) = self._hf_pts(W, d)
        uepc = self._hf_child_uep(W, d)
        ue2uc = [kernel(ucp, uepc + shifted_point(i, W), self.ktype) for i in range(8)]
        pd = parent_dir(d) if W != 1 else None
        for k in lists[0]:
            chk = np.zeros(len(ucp), dtype=np.complex128)
            for ind in range(8):
                ck = child_key(k, ind)
                if ck in boxes:
                    x = upeq[ck] if W == 1 else dirup[(ck, pd)]
                    chk = chk + translate(ue2uc[ind], x)
            dirup[(k, d)] = E1 @ (E2 @ (E3 @ chk))

    def _down_high(self, W, d, lists, boxes, dirup, dirdn, dnchk):
        kt = self.ktype
        uep, ucp, (E1, E2, E3) = self._hf_pts(W, d)
        dep, dcp = -ucp, -uep
        if W == 1:
            dcpc = self._lf(W / 2)[0]
        else:
            dcpc = -self._hf_pts(W / 2, parent_dir(d))[0]
        de2dc = [kernel(dcpc + shifted_point(i, W), dep, kt) for i in range(8)]
        pd = parent_dir(d) if W != 1 else None
        mcache = {}
        for k in lists[1]:
            b = boxes[k]
            c = self.center(k)
            dcv = dirdn.get((k, d), np.zeros(len(dcp), dtype=np.complex128))
            for sk in b.fnde.get(d, []):
                off = tuple(np.rint((c - self.center(sk)) / W).astype(np.int64))
                if off not in mcache:
                    mcache[off] = kernel(dcp + np.asarray(off, dtype=np.float64) * W, uep, kt)
                dcv = dcv + translate(mcache[off], dirup.get((sk, d), np.zeros(len(uep), dtype=np.complex128)))
            dirdn[(k, d)] = dcv
            dneq = E3.T @ (E2.T @ (E1.T @ dcv))
            for ind in range(8):
                ck = child_key(k, ind)
                if ck not in boxes:
                    continue
                if W == 1:
                    dnchk[ck] = dnchk.get(ck, 0) + translate(de2dc[ind], dneq)
                else:
                    dirdn[(ck, pd)] = dirdn.get((ck, pd), 0) + translate(de2dc[ind], dneq)


def dump_lists(w: "Wave3dOracle") -> str:
    """Canonical text of every non-empty box: tag, point count, fftnum, U/V/W/X/E lists,
    far lists by direction and the in/out direction sets (same layout as the C ABI's
    ddfmm_dump_lists)."""
    out = []

    def key(k):
        return f"{k[0]}:{k[1]},{k[2]},{k[3]}"

    def dr(d):
        return f"{d[0]},{d[1]},{d[2]}"

    for k, b in w.boxes.items():
        if not (b.tag & PTS):
            continue
        s = f"{key(k)} t{b.tag} n{len(b.ptidx)} f{b.fftnum}"
        for nm, lst in zip("UVWXE", (b.U, b.V, b.W, b.X, b.E)):
            s += f" {nm}[" + "".join(key(x) + " " for x in lst) + "]"
        for d in sorted(b.fnde):
            s += f" F({dr(d)})[" + "".join(key(x) + " " for x in b.fnde[d]) + "]"
        s += " I[" + "".join(dr(d) + " " for d in sorted(b.incdir)) + "]"
        s += " O[" + "".join(dr(d) + " " for d in sorted(b.outdir)) + "]"
        out.append(s + "\n")
    return "".join(out)
