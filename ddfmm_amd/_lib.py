"""ctypes binding of the C ABI in include/ddfmm.h (libddfmm.so, built in-tree)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddfmm.so")

EXPORTS = ["ddfmm_last_error", "ddfmm_create", "ddfmm_destroy", "ddfmm_setup", "ddfmm_eval",
           "ddfmm_eval_device", "ddfmm_direct_sum", "ddfmm_kernel_matrix", "ddfmm_get_stats",
           "ddfmm_dump_lists", "ddfmm_read_buffer", "ddfmm_set_timing", "ddfmm_stage_times",
           "ddfmm_stage_exec", "ddfmm_set_graph", "ddfmm_mlib_load", "ddfmm_mlib_free", "ddfmm_mlib_widths",
           "ddfmm_setup_mlib", "ddfmm_autotune", "ddfmm_set_tiles", "ddfmm_get_tiles", "ddfmm_direct_sum_noalloc",
           "ddfmm_set_hbm_budget", "ddfmm_memory"]


class Params(ctypes.Structure):
    _fields_ = [("K", ctypes.c_double), ("ctr", ctypes.c_double * 3), ("ACCU", ctypes.c_int),
                ("NPQ", ctypes.c_int), ("ptsmax", ctypes.c_int), ("maxlevel", ctypes.c_int),
                ("numC", ctypes.c_int), ("kernel", ctypes.c_int), ("nrhs_max", ctypes.c_int),
                ("precision", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "npts", "nleaf", "nlf", "nhf_out", "nhf_in", "near_interactions", "leaf_op_interactions",
        "s2m_interactions", "x2chk_interactions", "v_pairs", "v_targets", "v_sources", "m2l_pairs",
        "m2l_groups")] + [("m2l_flops", ctypes.c_double), ("gemm_flops", ctypes.c_double),
                          ("opmat_elems", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                          ("nlevels_lf", ctypes.c_int), ("nlevels_hf", ctypes.c_int),
                          ("max_hf_width", ctypes.c_int), ("nstages", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class DdfmmError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libddfmm.so; raises if the HIP library has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DdfmmError(f"{LIB_PATH} is missing: build it with ddfmm_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, ci, cd = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    L.ddfmm_last_error.restype = ctypes.c_char_p
    L.ddfmm_create.argtypes = [ci, ctypes.POINTER(vp)]
    L.ddfmm_destroy.argtypes = [vp]
    L.ddfmm_setup.argtypes = [vp, vp, i64, ctypes.POINTER(Params), ctypes.c_char_p, ctypes.c_char_p]
    L.ddfmm_eval.argtypes = [vp, vp, vp, ci]
    L.ddfmm_mlib_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.ddfmm_mlib_free.argtypes = [vp]
    L.ddfmm_mlib_widths.argtypes = [vp, ci, ctypes.POINTER(cd), ctypes.POINTER(ci), ctypes.POINTER(cd),
                                    ctypes.POINTER(ci)]
    L.ddfmm_setup_mlib.argtypes = [vp, vp, i64, ctypes.POINTER(Params), vp]
    L.ddfmm_eval_device.argtypes = [vp, vp, vp, ci, vp]
    L.ddfmm_direct_sum.argtypes = [ci, vp, ci, vp, ci, vp, ci, vp]
    L.ddfmm_direct_sum_noalloc.argtypes = [ci, vp, ci, vp, ci, vp, ci, vp, vp]
    L.ddfmm_kernel_matrix.argtypes = [ci, vp, ci, vp, ci, vp]
    L.ddfmm_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.ddfmm_dump_lists.argtypes = [vp, ctypes.c_char_p, i64, ctypes.POINTER(i64)]
    L.ddfmm_read_buffer.argtypes = [vp, ci, ci, i64, i64, vp]
    L.ddfmm_set_timing.argtypes = [vp, ci]
    L.ddfmm_set_graph.argtypes = [vp, ci]
    L.ddfmm_autotune.argtypes = [vp, ci, ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(cd),
                                 ctypes.POINTER(cd)]
    L.ddfmm_set_tiles.argtypes = [vp, ci, ctypes.POINTER(ci)]
    L.ddfmm_get_tiles.argtypes = [vp, ci, ctypes.POINTER(ci), ctypes.c_char_p, ctypes.POINTER(ci)]
    L.ddfmm_set_hbm_budget.argtypes = [vp, i64]
    L.ddfmm_memory.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(ci)]
    L.ddfmm_stage_times.argtypes = [vp, ci, ctypes.c_char_p, ctypes.POINTER(cd), ctypes.POINTER(cd),
                                     ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.ddfmm_stage_exec.argtypes = [vp, ci, ctypes.POINTER(cd), ctypes.POINTER(ci)]
    for name in EXPORTS:
        if name != "ddfmm_last_error":
            getattr(L, name).restype = ci
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise DdfmmError(lib().ddfmm_last_error().decode(errors="replace"))


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Hip:
    """Minimal HIP runtime calls (the same libamdhip64 that libddfmm.so links) for
    device-resident buffers and device-wide synchronisation in tests and benchmarks.
    torch is not used in processes that load libddfmm.so: its wheel bundles a different
    HIP runtime."""

    def __init__(self):
        lib()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [vp]
        self.rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.rt.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        self.rt.hipHostMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_uint]
        self.rt.hipHostFree.argtypes = [vp]
        self.rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.rt.hipStreamDestroy.argtypes = [vp]
        self.rt.hipStreamSynchronize.argtypes = [vp]
        self.rt.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        self.rt.hipEventDestroy.argtypes = [vp]
        self.rt.hipEventRecord.argtypes = [vp, vp]
        self.rt.hipEventQuery.argtypes = [vp]
        self.rt.hipEventSynchronize.argtypes = [vp]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipMemcpyAsync", "hipHostMalloc",
                  "hipHostFree", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize", "hipEventCreate",
                  "hipEventDestroy", "hipEventRecord", "hipEventQuery", "hipEventSynchronize"):
            getattr(self.rt, f).restype = ctypes.c_int

    def _ok(self, rc, what):
        if rc != 0:
            raise DdfmmError(f"{what} failed with HIP error {rc}")

    def malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, p: int):
        self._ok(self.rt.hipFree(p), "hipFree")

    def to_device(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def to_host(self, p: int, out: np.ndarray) -> np.ndarray:
        self._ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def copy_d2d(self, dst: int, src: int, nbytes: int):
        self._ok(self.rt.hipMemcpy(dst, src, nbytes, 3), "hipMemcpy D2D")

    def sync(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")

    # streams, events and pinned host memory (serving loop)
    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._ok(self.rt.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def stream_destroy(self, s: int):
        self._ok(self.rt.hipStreamDestroy(s), "hipStreamDestroy")

    def stream_sync(self, s: int):
        self._ok(self.rt.hipStreamSynchronize(s), "hipStreamSynchronize")

    def event(self) -> int:
        e = ctypes.c_void_p()
        self._ok(self.rt.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def event_destroy(self, e: int):
        self._ok(self.rt.hipEventDestroy(e), "hipEventDestroy")

    def record(self, e: int, s: int):
        self._ok(self.rt.hipEventRecord(e, s), "hipEventRecord")

    def done(self, e: int) -> bool:
        rc = self.rt.hipEventQuery(e)
        if rc == 0:
            return True
        if rc == 600:          # hipErrorNotReady
            return False
        raise DdfmmError(f"hipEventQuery failed with HIP error {rc}")

    def pinned(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._ok(self.rt.hipHostMalloc(ctypes.byref(p), nbytes, 0), "hipHostMalloc")
        return p.value

    def pinned_free(self, p: int):
        self._ok(self.rt.hipHostFree(p), "hipHostFree")

    def copy_async(self, dst: int, src: int, nbytes: int, kind: int, s: int):
        """kind: 1 host->device, 2 device->host, 3 device->device."""
        self._ok(self.rt.hipMemcpyAsync(dst, src, nbytes, kind, s), "hipMemcpyAsync")
