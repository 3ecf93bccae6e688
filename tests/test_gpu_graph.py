"""Graph replay (ddfmm_set_graph): the captured stage sequence gives the same bits as the
stage-by-stage evaluation, on every replay, with other device work and host
synchronisations between the replays, for several right-hand-side counts (one graph
each) and both precision modes; eval_device replays read the caller's buffers."""
import numpy as np
import pytest

from ddfmm_amd import geometry
from ddfmm_amd._lib import Hip
from ddfmm_amd.wave3d import FP64, MIXED, Mlib3d, Wave3d, direct_sum

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [MIXED, FP64])
def test_graph_replay_matches_eager(oplib, precision):
    K, N = 8.0, 20000
    pos = geometry.sphere(N, K, 31)
    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": oplib[0], "-mlib3d_hdname": oplib[1]})
    w = Wave3d()
    w.posptr, w.mlibptr = pos, ml
    w.setup({"-wave3d_ACCU": 2, "-wave3d_K": K, "-wave3d_maxlevel": 12}, nrhs_max=4, precision=precision)
    dens = {n: np.atleast_2d(geometry.densities(N, 40 + n, nrhs=n)) for n in (1, 3, 4)}
    eager = {n: w.eval(d) for n, d in dens.items()}
    w.set_graph(True)
    for rep in range(3):
        for n in (4, 1, 3):
            got = w.eval(dens[n])
            assert np.array_equal(got, eager[n]), (rep, n)
            # unrelated device work and a synchronisation between replays
            direct_sum(pos[:50], pos[:2000], dens[n][:, :2000])
    # device buffers through eval_device: the replay reads the pointers of this call
    hip = Hip()
    d = np.ascontiguousarray(dens[3])
    dd, vd = hip.to_device(d), hip.malloc(d.nbytes)
    dd2, vd2 = hip.to_device(np.ascontiguousarray(d[::-1])), hip.malloc(d.nbytes)
    try:
        w.eval_device(dd, vd, 3)
        w.eval_device(dd2, vd2, 3)
        hip.sync()
        a = hip.to_host(vd, np.empty_like(d))
        b = hip.to_host(vd2, np.empty_like(d))
        assert np.array_equal(a, eager[3])
        assert np.array_equal(b, eager[3][::-1])
    finally:
        for p in (dd, vd, dd2, vd2):
            hip.free(p)
    w.set_graph(False)
    assert np.array_equal(w.eval(dens[4]), eager[4])
    w.close()


def test_autotune_keeps_results(oplib):
    """The autotuner only changes how GEMM rows are cut into tiles; a column's k order is
    the same under every tile height, so the potentials keep their bits; the tuned
    table can be read back and applied to a fresh context of the same geometry."""
    K, N = 8.0, 30000
    pos = geometry.sphere(N, K, 32)
    den = geometry.densities(N, 33, nrhs=2)

    def make():
        ml = Mlib3d()
        ml.setup({"-mlib3d_ldname": oplib[0], "-mlib3d_hdname": oplib[1]})
        w = Wave3d()
        w.posptr, w.mlibptr = pos, ml
        w.setup({"-wave3d_ACCU": 2, "-wave3d_K": K, "-wave3d_maxlevel": 12}, nrhs_max=2, precision=MIXED)
        return w

    w = make()
    before = w.eval(den)
    res = w.autotune(nrhs=2, reps=1)
    table, names = w.tiles()
    print(f"autotune: {res['gemm_ms_planner']:.3f} -> {res['gemm_ms_tuned']:.3f} ms, tiles {table}", flush=True)
    assert table == res["tiles"] and len(names) == len(table) and all(1 <= m <= 7 for m in table)
    assert np.array_equal(w.eval(den), before)
    w.set_graph(True)
    assert np.array_equal(w.eval(den), before)
    w.close()
    w2 = make()
    w2.set_tiles(table)
    assert np.array_equal(w2.eval(den), before)
    w2.close()
