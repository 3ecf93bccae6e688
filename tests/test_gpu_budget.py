"""HBM budget (ddfmm_set_hbm_budget, the analogue of the reference's LIMITED_MEMORY
build, wave3d_eval.cpp:106-186): a context whose budget admits two right-hand sides per
pass evaluates five in three passes with the same bits as an unlimited context, through
the host-buffer and device-buffer entry points and under graph replay, and holds no more
device memory than the budget."""
import numpy as np
import pytest

from ddfmm_amd import geometry
from ddfmm_amd._lib import Hip
from ddfmm_amd.wave3d import FP64, MIXED, Mlib3d, Wave3d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [MIXED, FP64])
def test_budget_passes_match_one_pass(oplib, precision):
    K, N, R = 8.0, 30000, 5
    pos = geometry.sphere(N, K, 41)
    den = np.atleast_2d(geometry.densities(N, 42, nrhs=R))
    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": oplib[0], "-mlib3d_hdname": oplib[1]})

    def make(budget_gib):
        w = Wave3d()
        w.posptr, w.mlibptr = pos, ml
        w.setup({"-wave3d_ACCU": 2, "-wave3d_K": K, "-wave3d_maxlevel": 12, "-wave3d_hbm_budget": budget_gib},
                nrhs_max=R, precision=precision)
        return w

    full = make(0)
    ref = full.eval(den)
    m = full.memory()
    assert m["slab"] == R
    assert full.stats()["device_bytes"] == m["fixed_bytes"] + R * m["per_rhs_bytes"]
    full.close()
    budget = m["fixed_bytes"] + 2.5 * m["per_rhs_bytes"]
    w = make(budget / 2**30)
    mb = w.memory()
    used = w.stats()["device_bytes"]
    print(f"precision {precision}: tables {mb['fixed_bytes'] / 2**20:.1f} MiB, per rhs "
          f"{mb['per_rhs_bytes'] / 2**20:.1f} MiB, slab {mb['slab']}, held {used / 2**20:.1f} MiB", flush=True)
    assert mb["slab"] == 2 and used == mb["fixed_bytes"] + 2 * mb["per_rhs_bytes"] and used <= budget
    assert np.array_equal(w.eval(den), ref)
    assert np.array_equal(w.eval(den[:1]), ref[:1])
    hip = Hip()
    d = np.ascontiguousarray(den)
    dd, vd = hip.to_device(d), hip.malloc(d.nbytes)
    try:
        w.eval_device(dd, vd, R)
        hip.sync()
        assert np.array_equal(hip.to_host(vd, np.empty_like(d)), ref)
        w.set_graph(True)
        for _ in range(2):
            assert np.array_equal(w.eval(den), ref)
            w.eval_device(dd, vd, R)
            hip.sync()
            assert np.array_equal(hip.to_host(vd, np.empty_like(d)), ref)
    finally:
        hip.free(dd)
        hip.free(vd)
    w.close()
