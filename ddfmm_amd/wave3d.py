"""Python mirror of the reference's Wave3d / Kernel3d / Mlib3d surface over the C ABI.

Names, option keys and defaults follow include/wave3d.hpp:219-265,
include/kernel3d.hpp:29-45, include/mlib3d.hpp:58-99, the option parsing of
wave3d_setup.cpp:29-61 / mlib3d.cpp:70-91 and the driver wiring of tt.cpp:103-142.
Everything runs through libddfmm.so (HIP); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import Params, Stats, check, lib, ptr

KERNEL_HELM = 0
KERNEL_EXPR = 1
FP64 = 0          # FP64, solves factored as in the reference
MIXED = 1         # FP32 vectors/translations on MFMA, FP64 fused solves
FP64_FUSED = 2    # FP64 with fused solves
MIXED_SPLIT = 3   # MIXED with split-BF16 translation GEMMs


class Kernel3d:
    """Kernel3d (kernel3d.hpp:29-45): G = exp(ikr)/r (HELM) or exp(ikr) (EXPR), k = 2*pi."""

    def __init__(self, type: int = KERNEL_HELM):
        self.type = type

    def kernel(self, trgpos: np.ndarray, srcpos: np.ndarray, srcnor=None) -> np.ndarray:
        """Dense M x N block for 3 x M targets and 3 x N sources (column = point), as
        Kernel3d::kernel (kernel3d.cpp:26-107); srcnor is unused there too."""
        t = np.ascontiguousarray(np.asarray(trgpos, dtype=np.float64).T)
        s = np.ascontiguousarray(np.asarray(srcpos, dtype=np.float64).T)
        out = np.zeros((t.shape[0] * s.shape[0],), dtype=np.complex128)
        check(lib().ddfmm_kernel_matrix(self.type, ptr(t), t.shape[0], ptr(s), s.shape[0], ptr(out)))
        return out.reshape((t.shape[0], s.shape[0]), order="F")


class Mlib3d:
    """Operator-library handle (mlib3d.hpp:58-99): option keys -<prefix>NPQ/ldname/hdname.

    setup() reads the two files once into a library handle (ddfmm_mlib_load) that every
    Wave3d set up with this object shares; ``w2ldmap`` / ``w2hdmap`` list the widths of
    the LF and HF tables."""

    def __init__(self, prefix: str = "mlib3d_"):
        self.prefix = prefix
        self.kernel = Kernel3d()
        self.NPQ = 4
        self.ldname = ""
        self.hdname = ""
        self._h = None

    def setup(self, opts: dict) -> int:
        p = "-" + self.prefix
        if p + "NPQ" in opts:
            self.NPQ = int(opts[p + "NPQ"])
        self.ldname = str(opts.get(p + "ldname", self.ldname))
        self.hdname = str(opts.get(p + "hdname", self.hdname))
        self.close()
        h = ctypes.c_void_p()
        check(lib().ddfmm_mlib_load(self.ldname.encode(), self.hdname.encode(), ctypes.byref(h)))
        self._h = h
        return 0

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("Mlib3d.setup has not been called")
        return self._h

    def _widths(self):
        nld, nhd = ctypes.c_int(), ctypes.c_int()
        ld = (ctypes.c_double * 64)()
        hd = (ctypes.c_double * 64)()
        check(lib().ddfmm_mlib_widths(self.handle, 64, ld, ctypes.byref(nld), hd, ctypes.byref(nhd)))
        return list(ld[:nld.value]), list(hd[:nhd.value])

    @property
    def w2ldmap(self):
        return {w: None for w in self._widths()[0]}

    @property
    def w2hdmap(self):
        return {w: None for w in self._widths()[1]}

    def close(self):
        if self._h is not None:
            lib().ddfmm_mlib_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opt(opts, key, cast, default):
    return cast(opts[key]) if key in opts else default


class Wave3d:
    """Directional FMM evaluation (wave3d.hpp:219-265) on one MI355X.

    Wiring as tt.cpp:134-138: set ``posptr`` (N x 3 positions, row = point id),
    ``kernel``, ``mlibptr`` and ``geomprtn`` (cell partition; only its edge length
    numC matters on one device), then ``setup(opts)`` and ``eval(den)``.
    """

    def __init__(self, prefix: str = "wave3d_", device: int = 0):
        self.prefix = prefix
        self.device = device
        self.posptr = None
        self.kernel = Kernel3d()
        self.mlibptr = None
        self.geomprtn = np.zeros((1, 1, 1), dtype=np.int32)
        # defaults of Wave3d::Wave3d (wave3d.cpp:24-26)
        self.ACCU, self.NPQ, self.K = 1, 4, 64.0
        self.ctr = (0.0, 0.0, 0.0)
        self.ptsmax = 100
        self.maxlevel = None
        self.nrhs_max = 1
        self.precision = FP64
        self.hbm_budget = 0.0          # GiB; 0 = no budget (ddfmm_set_hbm_budget)
        self._ctx = None

    def setup(self, opts: dict | None = None, nrhs_max: int | None = None,
              precision: int | None = None) -> int:
        """Wave3d::setup (wave3d_setup.cpp:21-97)."""
        opts = opts or {}
        p = "-" + self.prefix
        self.ACCU = _opt(opts, p + "ACCU", int, self.ACCU)
        self.NPQ = _opt(opts, p + "NPQ", int, self.NPQ)
        self.K = _opt(opts, p + "K", float, self.K)
        if p + "ctr" in opts:
            self.ctr = tuple(float(x) for x in str(opts[p + "ctr"]).split())
        self.ptsmax = _opt(opts, p + "ptsmax", int, self.ptsmax)
        self.maxlevel = _opt(opts, p + "maxlevel", int, self.maxlevel)
        self.hbm_budget = _opt(opts, p + "hbm_budget", float, self.hbm_budget)
        if nrhs_max is not None:
            self.nrhs_max = int(nrhs_max)
        if precision is not None:
            self.precision = int(precision)
        if self.maxlevel is None:
            raise ValueError("maxlevel must be given (-wave3d_maxlevel); the reference leaves it uninitialised")
        if self.posptr is None or self.mlibptr is None:
            raise ValueError("posptr and mlibptr must be set before setup")
        pos = np.ascontiguousarray(self.posptr, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("posptr must be an N x 3 array")
        prm = Params(K=self.K, ACCU=self.ACCU, NPQ=self.NPQ, ptsmax=self.ptsmax,
                     maxlevel=self.maxlevel, numC=int(np.asarray(self.geomprtn).shape[0]),
                     kernel=self.kernel.type, nrhs_max=self.nrhs_max, precision=self.precision)
        prm.ctr[:] = self.ctr
        self.close()
        h = ctypes.c_void_p()
        check(lib().ddfmm_create(self.device, ctypes.byref(h)))
        self._ctx = h
        self._n = pos.shape[0]
        check(lib().ddfmm_set_hbm_budget(h, int(self.hbm_budget * 2**30)))
        check(lib().ddfmm_setup_mlib(h, ptr(pos), pos.shape[0], ctypes.byref(prm), self.mlibptr.handle))
        return 0

    def eval(self, den: np.ndarray) -> np.ndarray:
        """Wave3d::eval (wave3d_eval.cpp:299-415): den (N,) or (nrhs, N) complex."""
        d = np.ascontiguousarray(np.atleast_2d(np.asarray(den, dtype=np.complex128)))
        if d.shape[1] != self._n:
            raise ValueError("density length does not match the number of points")
        out = np.empty_like(d)
        check(lib().ddfmm_eval(self._ctx, ptr(d), ptr(out), d.shape[0]))
        return out[0] if np.ndim(den) == 1 else out

    def eval_device(self, den_ptr: int, val_ptr: int, nrhs: int, stream: int = 0) -> None:
        check(lib().ddfmm_eval_device(self._ctx, den_ptr, val_ptr, nrhs, stream))

    def check(self, den: np.ndarray, val: np.ndarray, chkkeys) -> float:
        """Wave3d::check (wave3d_check.cpp:22-100): ||direct - fmm|| / ||fmm|| at chkkeys."""
        keys = np.asarray(chkkeys, dtype=np.int64)
        pos = np.ascontiguousarray(self.posptr, dtype=np.float64)
        direct = direct_sum(pos[keys], pos, den, self.kernel.type)
        fmm = np.asarray(val)[..., keys]
        return float(np.linalg.norm(direct - fmm) / np.linalg.norm(fmm))

    def stats(self) -> dict:
        s = Stats()
        check(lib().ddfmm_get_stats(self._ctx, ctypes.byref(s)))
        return s.as_dict()

    def dump_lists(self) -> str:
        n = ctypes.c_int64()
        check(lib().ddfmm_dump_lists(self._ctx, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().ddfmm_dump_lists(self._ctx, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    def read_buffer(self, buf: int, off: int, count: int, rhs: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=np.complex128)
        check(lib().ddfmm_read_buffer(self._ctx, buf, rhs, off, count, ptr(out)))
        return out

    def set_timing(self, on=True):
        """Per-stage HIP-event timing of the following evaluations (ddfmm_set_timing);
        on = 2 issues every stage on one stream, so that each time is the stage's own."""
        check(lib().ddfmm_set_timing(self._ctx, int(on)))

    def autotune(self, nrhs: int = 1, reps: int = 2) -> dict:
        """Pick the tile-height limit of every GEMM stage by timing (ddfmm_autotune)."""
        cnt = ctypes.c_int()
        check(lib().ddfmm_get_tiles(self._ctx, 0, None, None, ctypes.byref(cnt)))
        n = cnt.value
        table = (ctypes.c_int * max(n, 1))()
        before, after = ctypes.c_double(), ctypes.c_double()
        check(lib().ddfmm_autotune(self._ctx, int(nrhs), int(reps), n, table, ctypes.byref(cnt),
                                   ctypes.byref(before), ctypes.byref(after)))
        return {"tiles": list(table[:n]), "gemm_ms_planner": before.value, "gemm_ms_tuned": after.value}

    def tiles(self) -> tuple[list, list]:
        """(per-stage tile-height limits, stage names) of the current GEMM tiling."""
        cnt = ctypes.c_int()
        check(lib().ddfmm_get_tiles(self._ctx, 0, None, None, ctypes.byref(cnt)))
        n = cnt.value
        table = (ctypes.c_int * max(n, 1))()
        names = ctypes.create_string_buffer(24 * max(n, 1))
        check(lib().ddfmm_get_tiles(self._ctx, n, table, names, ctypes.byref(cnt)))
        raw = names.raw
        return list(table[:n]), [raw[24 * i:24 * i + 24].split(b"\0")[0].decode() for i in range(n)]

    def set_tiles(self, table) -> None:
        t = (ctypes.c_int * max(len(table), 1))(*table)
        check(lib().ddfmm_set_tiles(self._ctx, len(table), t))

    def memory(self) -> dict:
        """Device bytes of the set-up context (ddfmm_memory): tables, one right-hand side's
        work buffers and staging, setup-only generator tables, and the right-hand sides
        per pass under the HBM budget."""
        f, r, t, s = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check(lib().ddfmm_memory(self._ctx, ctypes.byref(f), ctypes.byref(r), ctypes.byref(t), ctypes.byref(s)))
        return {"fixed_bytes": f.value, "per_rhs_bytes": r.value, "setup_bytes": t.value, "slab": s.value}

    def set_graph(self, on: bool = True):
        """Capture the evaluation of each right-hand-side count once into a HIP graph and
        replay it (ddfmm_set_graph)."""
        check(lib().ddfmm_set_graph(self._ctx, int(on)))

    def stage_times(self):
        """[(name, ms, work_per_rhs, kind, precision_class)] of the last timed eval."""
        cnt = ctypes.c_int()
        check(lib().ddfmm_stage_times(self._ctx, 0, None, None, None, None, ctypes.byref(cnt)))
        n = cnt.value
        names = ctypes.create_string_buffer(24 * n)
        ms = (ctypes.c_double * n)()
        work = (ctypes.c_double * n)()
        kinds = (ctypes.c_int * n)()
        check(lib().ddfmm_stage_times(self._ctx, n, names, ms, work, kinds, ctypes.byref(cnt)))
        return [(names.raw[24 * i:24 * i + 24].split(b"\0")[0].decode(), ms[i], work[i], kinds[i] % 16,
                 kinds[i] // 16) for i in range(n)]

    def stage_exec(self):
        """Per stage, the MFMA flops per RHS the GEMM kernel executes (0 for other stages)."""
        cnt = ctypes.c_int()
        check(lib().ddfmm_stage_exec(self._ctx, 0, None, ctypes.byref(cnt)))
        ex = (ctypes.c_double * cnt.value)()
        check(lib().ddfmm_stage_exec(self._ctx, cnt.value, ex, ctypes.byref(cnt)))
        return list(ex)

    def close(self):
        if self._ctx is not None:
            lib().ddfmm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def direct_sum(trg: np.ndarray, src: np.ndarray, den: np.ndarray, kernel: int = KERNEL_HELM) -> np.ndarray:
    """GPU FP64 direct sum (rows = points); den (n,) or (nrhs, n)."""
    t = np.ascontiguousarray(trg, dtype=np.float64)
    s = np.ascontiguousarray(src, dtype=np.float64)
    d = np.ascontiguousarray(np.atleast_2d(np.asarray(den, dtype=np.complex128)))
    out = np.zeros((d.shape[0], t.shape[0]), dtype=np.complex128)
    check(lib().ddfmm_direct_sum(kernel, ptr(t), t.shape[0], ptr(s), s.shape[0], ptr(d), d.shape[0], ptr(out)))
    return out[0] if np.ndim(den) == 1 else out


def plan_only(pos: np.ndarray, K: float, ldname: str, hdname: str, kernel: int = KERNEL_HELM,
              ACCU: int = 2, NPQ: int = 4, ptsmax: int = 100, maxlevel: int = 12, numC: int = 1):
    """Host-only planning context (no GPU): returns (stats dict, list dump text)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    prm = Params(K=K, ACCU=ACCU, NPQ=NPQ, ptsmax=ptsmax, maxlevel=maxlevel, numC=numC,
                 kernel=kernel, nrhs_max=1, precision=FP64)
    h = ctypes.c_void_p()
    check(lib().ddfmm_create(-1, ctypes.byref(h)))
    try:
        check(lib().ddfmm_setup(h, ptr(pos), pos.shape[0], ctypes.byref(prm), ldname.encode(), hdname.encode()))
        s = Stats()
        check(lib().ddfmm_get_stats(h, ctypes.byref(s)))
        n = ctypes.c_int64()
        check(lib().ddfmm_dump_lists(h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().ddfmm_dump_lists(h, buf, n.value, ctypes.byref(n)))
        return s.as_dict(), buf.raw[:n.value].decode()
    finally:
        lib().ddfmm_destroy(h)
