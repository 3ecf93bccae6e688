"""The compiled OpenMP oracle (oracle/wave3d_ref.cpp) against the numpy oracle and the
FP64 direct sum: the same tree and lists (box, leaf and near-interaction counts), the same
potentials to rounding for both kernels and on the adaptive geometry whose W and X lists
take their >= 152-point branches, and a planted mistake in the numpy oracle's V list that
the comparison catches."""
import numpy as np
import pytest

from ddfmm_amd import geometry
from oracle import wave3d_cref as C
from oracle import wave3d_ref as R


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def counts(ora):
    live = {k: b for k, b in ora.boxes.items() if b.tag & R.PTS}
    leaves = sum(1 for b in live.values() if b.tag & R.TERMINAL)
    near = sum(len(b.ptidx) * len(live[u].ptidx) for b in live.values() for u in b.U)
    lfb = sum(1 for k in live if ora.width(k) < 1 - 1e-12)
    return {"boxes": len(live), "leaves": leaves, "lf_boxes": lfb, "near": near}


@pytest.mark.parametrize("K,N,kt", [(4.0, 6000, 0), (8.0, 20000, 0), (4.0, 6000, 1)])
def test_compiled_matches_numpy_oracle(ops, oplib_expr, K, N, kt):
    from ddfmm_amd import serialize

    ld, hd = ops if kt == 0 else (serialize.read_ld(oplib_expr[0]), serialize.read_hd(oplib_expr[1]))
    pos = geometry.sphere(N, K, 5)
    den = geometry.densities(N, 5, nrhs=2)
    ora = R.Wave3dOracle(pos, K, ld, hd, ktype=kt).setup()
    ref = np.stack([ora.eval(den[r]) for r in range(2)])
    c = C.Wave3dCOracle(pos, K, ld, hd, ktype=kt).setup()
    assert c.stats() == counts(ora)
    got = c.eval(den, threads=4)
    tol = 1e-13 if kt == 0 else 1e-11      # e^{ikr}: larger equivalent densities, same rounding
    assert rel(got, ref) < tol, rel(got, ref)
    chk = geometry.check_targets(N, 200, 5)
    exact = R.direct_sum(pos[chk], pos, den, kt)
    if kt == 0:
        assert rel(got[:, chk], exact) < 1e-5
    c.close()


def test_compiled_adaptive_w_x_branches(ops):
    a = geometry.sphere(20000, 3.4, 1)
    b = geometry.sphere(80000, 3.4, 2)
    pos, K = np.concatenate([a, b[b[:, 2] > 1.2]]), 4.0
    den = geometry.densities(len(pos), 9)
    ora = R.Wave3dOracle(pos, K, *ops, ptsmax=400).setup()
    ref = ora.eval(den)
    c = C.Wave3dCOracle(pos, K, *ops, ptsmax=400).setup()
    assert c.stats() == counts(ora)
    assert rel(c.eval(den), ref) < 1e-13
    c.close()


def test_planted_vlist_mistake_is_caught(ops):
    """Dropping one V-list entry in the numpy oracle moves its potentials by far more than
    the rounding-level agreement the comparison demands."""
    K, N = 4.0, 6000
    pos = geometry.sphere(N, K, 6)
    den = geometry.densities(N, 6)
    c = C.Wave3dCOracle(pos, K, *ops).setup()
    good = c.eval(den)
    ora = R.Wave3dOracle(pos, K, *ops).setup()
    victim = next(b for b in ora.boxes.values() if len(b.V) > 1)
    victim.V = victim.V[1:]
    assert rel(ora.eval(den), good) > 1e-8
    c.close()


def test_bench_cpu_baseline_leg(oplib):
    """bench.py's CPU-baseline leg: the compiled oracle on a workload, one RHS, eval timed,
    the thread count reported and the result checked against the direct sum."""
    import bench

    out = bench.cpu_baseline(oplib, 0, "sphere8_60k")
    assert out["kind"] == "port" and 1 <= out["cores"] <= 16 and out["value"] > 0
    assert "sphere8_60k" in out["sample"] and out["eval_s"] > 0
    err = float(out["sample"].split("rel.err ")[1].split(" ")[0])
    assert err < 1e-5
