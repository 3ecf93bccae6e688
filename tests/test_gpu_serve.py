"""Serving on the GPU: continuous batching over two resident geometries; every result is
bit-for-bit the same alone and in the mix (each right-hand side is computed on its own
with a fixed summation order), and all results stay within 1e-5 of the FP64 direct sum's
accuracy bar."""
import numpy as np
import pytest

from ddfmm_amd import geometry
from ddfmm_amd.serve import GpuEngine, Request, Server, latency_stats, poisson_requests
from ddfmm_amd.wave3d import MIXED

pytestmark = pytest.mark.gpu

GEOMS = {"s4": (4.0, 5000, 11), "s8": (8.0, 20000, 12)}


def _engines(oplib, slot=8):
    pts = {k: geometry.sphere(n, K, seed) for k, (K, n, seed) in GEOMS.items()}

    def make(key):
        K = GEOMS[key][0]
        return GpuEngine(key, pts[key], {"-wave3d_ACCU": 2, "-wave3d_K": K, "-wave3d_maxlevel": 12},
                         oplib, slot, MIXED)

    return pts, make


def test_alone_equals_in_mix(oplib):
    pts, make = _engines(oplib)
    srv = Server(make, lambda key: 1 << 30, hbm_budget=64 << 30)
    reqs = poisson_requests({k: v[1] for k, v in GEOMS.items()}, qps=400, count=24, seed=5, max_rhs=4)
    srv.run(reqs)
    st = latency_stats(reqs)
    print(f"served {st['served']} batches {srv.batches} p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms",
          flush=True)
    assert all(r.error is None for r in reqs)
    assert srv.batches < len(reqs)                 # requests were batched together
    # the same requests served one at a time on the same resident engines
    for r in reqs[:10]:
        alone = Request(1000 + r.rid, r.geometry, r.den, 0.0)
        srv.run([alone])
        assert np.array_equal(alone.result, r.result), r.rid
    srv.close()


def test_eviction_under_budget(oplib):
    pts, make = _engines(oplib, slot=4)
    probe = make("s8")
    b = probe.bytes
    probe.close()
    srv = Server(make, lambda key: b, hbm_budget=int(1.5 * b))   # only one engine resident at a time
    reqs = poisson_requests({k: v[1] for k, v in GEOMS.items()}, qps=200, count=8, seed=6, max_rhs=2)
    srv.run(reqs)
    assert all(r.error is None for r in reqs)
    assert srv.evictions >= 1
    srv.close()


def test_tt_driver_reference_formats(tmp_path):
    """tt-style driver on files in the reference layout: RESULT block, Ea within 1e-5 and the
    written val file equal to a direct evaluation through the API."""
    import io

    from ddfmm_amd import serialize, tt

    root = tmp_path / "data"
    pos, den, ld, hd = tt.make_data(str(root), "sphere8_60k", nchk=300)
    buf = io.StringIO()
    res = tt.run({"-dataroot": str(root), "-posfile": "pos", "-denfile": "den", "-geomprtn": "geomprtn",
                  "-chkfile": "chk", "-valfile": "val", "-kernel": "0", "-mlib3d_ldname": ld,
                  "-mlib3d_hdname": hd, "-wave3d_ACCU": "2", "-wave3d_NPQ": "4", "-wave3d_K": "8",
                  "-wave3d_ptsmax": "100", "-wave3d_maxlevel": "12"}, out=buf)
    text = buf.getvalue()
    print(text, flush=True)
    assert "RESULT" in text and res["Ea"] < 1e-5
    keys, val = serialize.read_parvec_cpx(str(root / "val_0_1"))
    assert np.array_equal(keys, np.arange(len(pos)))
    from oracle import wave3d_ref as R
    chk = geometry.check_targets(len(pos), 200, 0)
    exact = R.direct_sum(pos[chk], pos, den)
    assert np.linalg.norm(val[chk] - exact) / np.linalg.norm(exact) < 1e-5


def test_config4_serving_mix_real_geometries(oplib):
    """The serving mix over the four benchmark geometries (60k-4M points) resident at once:
    requests of 1-4 RHS batched per geometry; one request per geometry replayed alone gives
    the same bits as in the mix, and each result is within 1e-5 of the direct sum."""
    from ddfmm_amd.wave3d import direct_sum

    names = ["sphere8_60k", "sphere16_250k_x64", "sphere32_1m", "ellipsoid_4m_x8"]
    pts = {n: geometry.workload_points(n) for n in names}

    def make(key):
        wl = geometry.WORKLOADS[key]
        return GpuEngine(key, pts[key], {"-wave3d_ACCU": 2, "-wave3d_K": wl["K"], "-wave3d_ptsmax": wl["ptsmax"],
                                         "-wave3d_maxlevel": 14}, oplib, 4, MIXED, numC=wl.get("numC", 1))

    srv = Server(make, lambda key: 0, hbm_budget=200 << 30)
    reqs = poisson_requests({n: len(pts[n]) for n in names}, qps=50, count=16, seed=21, max_rhs=4)
    srv.run(reqs)
    st = latency_stats(reqs)
    print(f"mix: served {st['served']} in {srv.batches} batches, p50 {st['p50_ms']:.1f} ms "
          f"p99 {st['p99_ms']:.1f} ms", flush=True)
    assert all(r.error is None for r in reqs)
    seen = set()
    for r in reqs:
        if r.geometry in seen:
            continue
        seen.add(r.geometry)
        alone = Request(2000 + r.rid, r.geometry, r.den, 0.0)
        srv.run([alone])
        assert np.array_equal(alone.result, r.result), (r.geometry, r.rid)
        P = pts[r.geometry]
        chk = geometry.check_targets(len(P), 300, 21)
        exact = direct_sum(P[chk], P, np.atleast_2d(r.den))
        e = float(np.linalg.norm(np.atleast_2d(r.result)[:, chk] - exact) / np.linalg.norm(exact))
        print(f"  {r.geometry} rid {r.rid} nrhs {r.nrhs}: alone == in mix, rel.err {e:.2e}", flush=True)
        assert e < 1e-5
    assert len(seen) == 4
    srv.close()
