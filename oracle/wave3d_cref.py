"""ORACLE (test infrastructure only) — ctypes front of ``libwave3d_ref.so``, the compiled
OpenMP restatement of the reference's directional FMM (oracle/wave3d_ref.cpp).

Only tests and the CPU-baseline leg of ``bench.py`` use it: as a full-size FP64 checker
and as the timed CPU path.  Same constructor arguments as ``wave3d_ref.Wave3dOracle``;
the operator library comes from the reference-format files through
``ddfmm_amd.serialize.read_ld`` / ``read_hd`` (point sets stored 3 x n, matrices as
complex arrays, V-list tensors keyed by offset).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libwave3d_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libwave3d_ref.so missing: run `make -C oracle`")
        # idle OpenMP workers sleep instead of spinning, so that numpy's BLAS threads in
        # the same process (the numpy oracle, emulations) keep their cores
        os.environ.setdefault("OMP_WAIT_POLICY", "passive")
        L = ctypes.CDLL(path)
        vp, ci, cd, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        L.wr_error.restype = ctypes.c_char_p
        L.wr_create.restype = vp
        L.wr_create.argtypes = [vp, i64, cd, vp, ci, ci, ci, ci, ci, ci]
        L.wr_free.argtypes = [vp]
        L.wr_add_lf.argtypes = [vp, cd, ci, vp, ci, vp, ci, ci, vp, vp, ci, ci, vp, ci, vp, vp]
        L.wr_add_hf.argtypes = [vp, cd, vp, ci, vp, ci, vp, ci, ci, vp, ci, ci, vp, ci, ci, vp]
        L.wr_setup.argtypes = [vp]
        L.wr_stats.argtypes = [vp, vp]
        L.wr_eval.argtypes = [vp, vp, vp, ci]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().wr_error().decode())


class Wave3dCOracle:
    """Wave3d::setup + eval (FP64, one right-hand side per pass) on the host cores."""

    def __init__(self, pos, K, ld, hd, ktype=0, ACCU=2, NPQ=4, ctr=(0.0, 0.0, 0.0), ptsmax=100,
                 maxlevel=12, numC=1):
        self.pos = _c(pos, np.float64)
        self.N = len(self.pos)
        self.P = {1: 4, 2: 6, 3: 8}[ACCU]
        ctr = _c(ctr, np.float64)
        L = lib()
        self._h = L.wr_create(self.pos.ctypes.data, self.N, float(K), ctr.ctypes.data, int(ktype), self.P,
                              int(NPQ), int(ptsmax), int(maxlevel), int(numC))
        keep = []
        for W, e in ld.items():
            uep, ucp = _c(e["uep"].T, np.float64), _c(e["ucp"].T, np.float64)
            V, invs, UH = (_c(m, np.complex128) for m in e["uc2ue"])
            offs = sorted(e["ue2dc"])
            o = _c(np.array(offs, dtype=np.int32).reshape(-1, 3), np.int32)
            t = _c(np.stack([e["ue2dc"][k] for k in offs]) if offs else np.zeros((0,)), np.complex128)
            keep += [uep, ucp, V, invs, UH, o, t]
            _check(L.wr_add_lf(self._h, float(W), len(uep), uep.ctypes.data, len(ucp), ucp.ctypes.data,
                               V.shape[0], V.shape[1], V.ctypes.data, invs.ctypes.data, UH.shape[0], UH.shape[1],
                               UH.ctypes.data, len(offs), o.ctypes.data, t.ctypes.data))
        for W, inner in hd.items():
            for srt, e in inner.items():
                uep, ucp = _c(e["uep"].T, np.float64), _c(e["ucp"].T, np.float64)
                E = [_c(m, np.complex128) for m in e["uc2ue"]]
                s = _c(np.array(srt, dtype=np.int32), np.int32)
                keep += [uep, ucp, s] + E
                _check(L.wr_add_hf(self._h, float(W), s.ctypes.data, len(uep), uep.ctypes.data, len(ucp),
                                   ucp.ctypes.data, E[0].shape[0], E[0].shape[1], E[0].ctypes.data,
                                   E[1].shape[0], E[1].shape[1], E[1].ctypes.data, E[2].shape[0],
                                   E[2].shape[1], E[2].ctypes.data))
        del keep

    def setup(self):
        _check(lib().wr_setup(self._h))
        return self

    def stats(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        _check(lib().wr_stats(self._h, out.ctypes.data))
        return {"boxes": int(out[0]), "leaves": int(out[1]), "lf_boxes": int(out[2]), "near": int(out[3])}

    def eval(self, den: np.ndarray, threads: int = 0) -> np.ndarray:
        """den (N,) or (nrhs, N) complex; one pass per right-hand side."""
        d2 = _c(np.atleast_2d(den), np.complex128)
        out = np.empty_like(d2)
        for r in range(d2.shape[0]):
            _check(lib().wr_eval(self._h, d2[r].ctypes.data, out[r].ctypes.data, int(threads)))
        return out[0] if np.ndim(den) == 1 else out

    def close(self):
        if self._h:
            lib().wr_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
