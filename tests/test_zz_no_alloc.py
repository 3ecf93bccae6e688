"""The no-allocation direct sum (ddfmm_direct_sum_noalloc, device buffers, partial sums in
the library's static array) against the allocating entry point ddfmm_direct_sum on the
same inputs: bit-identical, on two calls in a row, the second one smaller, so the second
call reuses the static array the first one filled."""
import ctypes

import numpy as np
import pytest

from ddfmm_amd import _lib
from ddfmm_amd._lib import Hip, check, lib

pytestmark = pytest.mark.gpu


def _direct_host(kernel, trg, src, den):
    nrhs = den.shape[0]
    out = np.zeros((nrhs, len(trg)), dtype=np.complex128)
    check(lib().ddfmm_direct_sum(kernel, trg.ctypes.data, len(trg), src.ctypes.data, len(src),
                                 den.ctypes.data, nrhs, out.ctypes.data))
    return out


def _direct_noalloc(hip, kernel, trg, src, den):
    nrhs = den.shape[0]
    dt, ds, dd = hip.to_device(trg), hip.to_device(src), hip.to_device(den)
    out = np.zeros((nrhs, len(trg)), dtype=np.complex128)
    do = hip.malloc(out.nbytes)
    try:
        check(lib().ddfmm_direct_sum_noalloc(kernel, dt, len(trg), ds, len(src), dd, nrhs, do, None))
        hip.sync()
        hip.to_host(do, out)
    finally:
        for p in (dt, ds, dd, do):
            hip.free(p)
    return out


@pytest.mark.parametrize("kernel", [0, 1])
def test_noalloc_direct_sum_matches_allocating_entry(kernel):
    hip = Hip()
    rng = np.random.default_rng(40 + kernel)
    # first call: 1000 targets x 200k sources x 8 RHS (many source ranges, all of the
    # partial-sum array in use); second call smaller: 300 targets x 50k sources x 3 RHS
    for m, n, nrhs in ((1000, 200_000, 8), (300, 50_000, 3), (70, 3000, 1)):
        src = rng.uniform(-4, 4, (n, 3))
        trg = np.ascontiguousarray(src[rng.choice(n, m, replace=False)])   # includes coincident pairs
        den = np.ascontiguousarray(rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n)))
        ref = _direct_host(kernel, trg, src, den)
        got = _direct_noalloc(hip, kernel, trg, src, den)
        assert np.array_equal(got, ref), (m, n, nrhs, float(np.abs(got - ref).max()))


def test_noalloc_rejects_bad_arguments():
    with pytest.raises(_lib.DdfmmError, match="bad argument"):
        check(lib().ddfmm_direct_sum_noalloc(0, None, 10, None, 10, None, 1, None, None))
    with pytest.raises(_lib.DdfmmError, match="Unknown kernel"):
        p = ctypes.c_void_p(16)
        check(lib().ddfmm_direct_sum_noalloc(3, p, 10, p, 10, p, 1, p, None))
