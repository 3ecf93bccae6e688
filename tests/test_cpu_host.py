"""CPU checks of the C ABI library: it loads, exports every declared symbol, and its
host planner (tree, U/V/W/X lists, far lists, direction sets) is integer-identical to
the oracle's restatement of wave3d_setup.cpp."""
import ctypes
import os
import re

import numpy as np
import pytest

from ddfmm_amd import _lib, geometry
from ddfmm_amd.wave3d import plan_only
from oracle import wave3d_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "ddfmm.h")).read()
    names = set(re.findall(r"\b(ddfmm_[a-z_]+)\s*\(", hdr))
    assert names, "no declarations found"
    L = _lib.lib()
    for n in sorted(names):
        assert hasattr(L, n), f"libddfmm.so does not export {n}"
    assert names == set(_lib.EXPORTS)


@pytest.mark.parametrize("K,N,seed,ptsmax,numC", [(4.0, 6000, 0, 100, 1), (8.0, 60000, 0, 100, 1),
                                                   (8.0, 20000, 5, 30, 2)])
def test_lists_match_oracle(oplib, ops, K, N, seed, ptsmax, numC):
    pos = geometry.sphere(N, K, seed)
    st, dump = plan_only(pos, K, *oplib, ptsmax=ptsmax, numC=numC)
    w = R.Wave3dOracle(pos, K, *ops, ptsmax=ptsmax, numC=numC).setup()
    assert dump == R.dump_lists(w)
    # statistics derived from the same lists
    live = {k: b for k, b in w.boxes.items() if b.tag & R.PTS}
    leaves = [k for k, b in live.items() if b.tag & R.TERMINAL]
    assert st["nleaf"] == len(leaves)
    assert st["npts"] == N
    assert st["v_pairs"] == sum(len(b.V) for b in live.values())


def test_adaptive_ellipsoid_lists(oplib, ops):
    pos = geometry.ellipsoid(30000, axes=(3.6, 1.0, 1.0), eps=0.1, seed=3)
    K = 8.0
    st, dump = plan_only(pos, K, *oplib, ptsmax=40)
    w = R.Wave3dOracle(pos, K, *ops, ptsmax=40).setup()
    assert dump == R.dump_lists(w)
    assert st["x2chk_interactions"] + st["near_interactions"] > 0


def test_errors_are_reported(oplib, tmp_path):
    pos = geometry.sphere(100, 4.0, 0)
    with pytest.raises(_lib.DdfmmError, match="outside"):
        plan_only(pos * 3, 4.0, *oplib)
    with pytest.raises(_lib.DdfmmError, match="Unknown kernel"):
        plan_only(pos, 4.0, *oplib, kernel=2)
    with pytest.raises(_lib.DdfmmError, match="power of two"):
        plan_only(pos, 6.0, *oplib)
    with pytest.raises(_lib.DdfmmError, match="cannot open"):
        plan_only(pos, 4.0, str(tmp_path / "missing.bin"), oplib[1])
    # an operator library without the needed HF width is rejected with a clear message
    from ddfmm_amd import serialize
    hd = serialize.read_hd(oplib[1])
    serialize.write_hd(tmp_path / "hd1.bin", {1.0: hd[1.0]})
    with pytest.raises(_lib.DdfmmError, match="high-frequency width 2"):
        plan_only(geometry.sphere(20000, 8.0, 0), 8.0, oplib[0], str(tmp_path / "hd1.bin"))


def test_eval_without_device_is_rejected(oplib):
    pos = geometry.sphere(500, 4.0, 0)
    prm = _lib.Params(K=4.0, ACCU=2, NPQ=4, ptsmax=100, maxlevel=12, numC=1, kernel=0, nrhs_max=1, precision=0)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ddfmm_create(-1, ctypes.byref(h)))
    try:
        _lib.check(_lib.lib().ddfmm_setup(h, pos.ctypes.data, 500, ctypes.byref(prm), oplib[0].encode(), oplib[1].encode()))
        den = np.zeros(500, complex)
        with pytest.raises(_lib.DdfmmError, match="not set up on a device"):
            _lib.check(_lib.lib().ddfmm_eval(h, den.ctypes.data, den.ctypes.data, 1))
    finally:
        _lib.lib().ddfmm_destroy(h)


def _host_stages(oplib, pos, K, precision, nrhs_max=1, ptsmax=100, dump=False):
    prm = _lib.Params(K=K, ACCU=2, NPQ=4, ptsmax=ptsmax, maxlevel=12, numC=1, kernel=0, nrhs_max=nrhs_max,
                      precision=precision)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ddfmm_create(-1, ctypes.byref(h)))
    try:
        _lib.check(_lib.lib().ddfmm_setup(h, pos.ctypes.data, len(pos), ctypes.byref(prm), oplib[0].encode(),
                                          oplib[1].encode()))
        cnt = ctypes.c_int()
        _lib.check(_lib.lib().ddfmm_stage_times(h, 0, None, None, None, None, ctypes.byref(cnt)))
        n = cnt.value
        names = ctypes.create_string_buffer(24 * n)
        work = (ctypes.c_double * n)()
        kinds = (ctypes.c_int * n)()
        _lib.check(_lib.lib().ddfmm_stage_times(h, n, names, None, work, kinds, ctypes.byref(cnt)))
        st = _lib.Stats()
        _lib.check(_lib.lib().ddfmm_get_stats(h, ctypes.byref(st)))
        out = ([(names.raw[24 * i:24 * i + 24].split(b"\0")[0].decode(), work[i], kinds[i] % 16, kinds[i] // 16)
                for i in range(n)], st.as_dict())
        if dump:
            ln = ctypes.c_int64()
            _lib.check(_lib.lib().ddfmm_dump_lists(h, None, 0, ctypes.byref(ln)))
            buf = ctypes.create_string_buffer(ln.value + 1)
            _lib.check(_lib.lib().ddfmm_dump_lists(h, buf, ln.value, ctypes.byref(ln)))
            out += (buf.raw[:ln.value].decode(),)
        return out
    finally:
        _lib.lib().ddfmm_destroy(h)


@pytest.mark.parametrize("precision", [0, 1])
def test_stage_precision_classes_and_work(oplib, precision):
    """Solve stages (check -> equivalent: LF pinv, HF E1 E2 E3 chains) carry the solve class
    that the mixed mode runs in FP64; every GEMM launch (one per tile height) keeps it; the
    per-launch useful flops add up to the planner's GEMM total.  In the mixed mode the HF
    solves are folded into the translations at setup, so no HF solve stage remains."""
    pos = geometry.sphere(20000, 8.0, 0)
    stages, st = _host_stages(oplib, pos, 8.0, precision)
    ST_GEMM, ST_REDUCE = 4, 7
    solve_names = ("uc2ue", "dc2de", "hf_up_e", "hf_dn_e")
    gemm = [s for s in stages if s[2] == ST_GEMM]
    assert gemm and any(s[2] == ST_REDUCE for s in stages)
    for name, _, kind, prec in stages:
        assert prec == (1 if any(t in name for t in solve_names) else 0), name
    names = {s[0] for s in gemm}
    if precision == 1:
        assert {"lf_uc2ue", "lf_dc2de", "hf_u2u", "hf_l2l"} <= names
        assert not any(n.startswith("hf_up_e") or n.startswith("hf_dn_e") for n in names)
    else:
        assert {"lf_uc2ue_1", "lf_uc2ue_2", "hf_up_e3", "hf_dn_e1t"} <= names
    total = sum(s[1] for s in gemm)
    assert abs(total - st["gemm_flops"]) <= 1e-9 * st["gemm_flops"]


def test_near_pairs_cover_every_interaction(oplib):
    """Mixed contexts on the matrix-core kernel apply (4 or more RHS, leaves of at most 128
    points) evaluate a near-field pair of leaves once for both directions when each is in
    the other's near list (U, W-direct, X-direct) and the smaller has at most 64 points.
    Checked against the list dump: the near-field launches still count every interaction
    once per direction, the mirror reduction adds one slot per such pair, of the larger
    leaf's size, and the one-way plan (fewer RHS) has neither."""
    a = geometry.sphere(20000, 3.4, 1)
    b = geometry.sphere(80000, 3.4, 2)
    pos = np.ascontiguousarray(np.concatenate([a, b[b[:, 2] > 1.2]]))
    one, st1 = _host_stages(oplib, pos, 4.0, 1, nrhs_max=1)
    two, st2, dump = _host_stages(oplib, pos, 4.0, 1, nrhs_max=8, dump=True)
    ST_KAPPLY, ST_REDUCE = 2, 7
    work = lambda stages, names: sum(w for n, w, k, p in stages if n in names)
    names1, names2 = [s[0] for s in one], [s[0] for s in two]
    assert "leaf_near_pairs" not in names1 and "near_mirror" not in names1
    assert "leaf_near_pairs" in names2 and "near_mirror" in names2
    assert st1["near_interactions"] == st2["near_interactions"] > 0
    assert work(one, {"leaf_near"}) == st1["near_interactions"]
    assert work(two, {"leaf_near", "leaf_near_pairs"}) == st2["near_interactions"]
    for other in ("leaf_far", "lf_s2m", "lf_x2chk"):
        assert work(one, {other}) == work(two, {other}), other
    assert [s[2] for s in two if s[0] == "near_mirror"] == [ST_REDUCE]
    # the pairs from the list dump: boxes (tag bit 2 terminal, points), their near lists
    NL = 152
    box = {}
    for ln in dump.splitlines():
        m = re.match(r"^(\S+) t(\d+) n(\d+) f\d+ U\[([^\]]*)\] V\[[^\]]*\] W\[([^\]]*)\] X\[([^\]]*)\]", ln)
        if m:
            box[m[1]] = (int(m[2]) & 2 != 0, int(m[3]), m[4].split(), m[5].split(), m[6].split())
    cnt = {}
    for k, (term, n, U, W, X) in box.items():
        if not term or n == 0:
            continue
        near = list(U) + [s for s in W if box[s][0] and box[s][1] < NL] + (list(X) if n < NL else [])
        for s in near:
            cnt[(k, s)] = cnt.get((k, s), 0) + 1
    expect = 0
    for (k, s), c in cnt.items():
        if k < s and c == 1 and cnt.get((s, k)) == 1:
            nk, ns = box[k][1], box[s][1]
            if min(nk, ns) <= 64:
                expect += max(nk, ns)
    assert expect > 0
    assert work(two, {"near_mirror"}) == expect


def test_blocked_gemm_padding_stays_small(oplib):
    """Column packing of the K-blocked GEMMs (child masks in U2U, direction bundles in the
    width-1 L2L): on the 250k-point sphere the largest U2U and L2L launches execute at most
    30% more MFMA work than the useful flops (a plain Gray-code order gave 1.3-1.6 at
    width 2), and the width-1 L2L bundles leave the LF reduction a few slots per child."""
    pos = geometry.sphere(250_000, 16.0, 2)
    prm = _lib.Params(K=16.0, ACCU=2, NPQ=4, ptsmax=100, maxlevel=12, numC=1, kernel=0, nrhs_max=1,
                      precision=1)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ddfmm_create(-1, ctypes.byref(h)))
    try:
        _lib.check(_lib.lib().ddfmm_setup(h, pos.ctypes.data, len(pos), ctypes.byref(prm), oplib[0].encode(),
                                          oplib[1].encode()))
        cnt = ctypes.c_int()
        _lib.check(_lib.lib().ddfmm_stage_times(h, 0, None, None, None, None, ctypes.byref(cnt)))
        n = cnt.value
        names = ctypes.create_string_buffer(24 * n)
        work = (ctypes.c_double * n)()
        kinds = (ctypes.c_int * n)()
        ex = (ctypes.c_double * n)()
        _lib.check(_lib.lib().ddfmm_stage_times(h, n, names, None, work, kinds, ctypes.byref(cnt)))
        _lib.check(_lib.lib().ddfmm_stage_exec(h, n, ex, ctypes.byref(cnt)))
        st = [(names.raw[24 * i:24 * i + 24].split(b"\0")[0].decode(), work[i], ex[i], kinds[i] % 16)
              for i in range(n)]
    finally:
        _lib.lib().ddfmm_destroy(h)
    for name in ("hf_u2u", "hf_l2l"):
        big = max((s for s in st if s[0] == name), key=lambda s: s[1])
        assert big[1] > 0 and big[2] / big[1] <= 1.3, (name, big[2] / big[1])
    # width-1 L2L: useful flops 8 * 152 * kb per (LF child, present direction), one slot of
    # 152 elements per (LF child, bundle); kb (the width-1 solve size) is at most 128, so
    # flops / (8 * 128) / slot elements bounds the mean number of directions per slot
    last_red = max(i for i, s in enumerate(st) if s[0] == "hf_reduce")      # width 1 comes last
    l2l_w1 = sum(s[1] for s in st[last_red:] if s[0] == "hf_l2l")
    lf_red = sum(s[1] for s in st if s[0] == "lf_reduce")
    assert lf_red > 0 and l2l_w1 / (8.0 * 128) / lf_red >= 3.0


def test_retile_keeps_useful_work(oplib):
    """Rebuilding the GEMM tiles with another tile-height limit (the autotuner's search
    space) changes the launches and the executed MFMA work, never the useful work."""
    from ddfmm_amd._lib import Stats

    pos = geometry.sphere(6000, 4.0, 0)
    prm = _lib.Params(K=4.0, ACCU=2, NPQ=4, ptsmax=100, maxlevel=12, numC=1, kernel=0, nrhs_max=1, precision=1)
    h = ctypes.c_void_p()
    L = _lib.lib()
    _lib.check(L.ddfmm_create(-1, ctypes.byref(h)))
    try:
        _lib.check(L.ddfmm_setup(h, pos.ctypes.data, len(pos), ctypes.byref(prm), oplib[0].encode(),
                                 oplib[1].encode()))

        def work():
            n = ctypes.c_int()
            _lib.check(L.ddfmm_stage_times(h, 0, None, None, None, None, ctypes.byref(n)))
            w = (ctypes.c_double * n.value)()
            k = (ctypes.c_int * n.value)()
            ex = (ctypes.c_double * n.value)()
            _lib.check(L.ddfmm_stage_times(h, n.value, None, None, w, k, ctypes.byref(n)))
            _lib.check(L.ddfmm_stage_exec(h, n.value, ex, ctypes.byref(n)))
            gemm = [i for i in range(n.value) if k[i] % 16 == 4]
            return sum(w[i] for i in gemm), sum(ex[i] for i in gemm), len(gemm)

        cnt = ctypes.c_int()
        _lib.check(L.ddfmm_get_tiles(h, 0, None, None, ctypes.byref(cnt)))
        nl = cnt.value
        assert nl >= 3
        u0, e0, n0 = work()
        for m in (1, 3, 7):
            tab = (ctypes.c_int * nl)(*([m] * nl))
            _lib.check(L.ddfmm_set_tiles(h, nl, tab))
            got = (ctypes.c_int * nl)()
            _lib.check(L.ddfmm_get_tiles(h, nl, got, None, ctypes.byref(cnt)))
            assert list(got) == [m] * nl
            u, e, n = work()
            assert u == pytest.approx(u0, rel=1e-12)
            assert e >= u
        bad = (ctypes.c_int * nl)(*([8] * nl))
        with pytest.raises(_lib.DdfmmError, match="out of range"):
            _lib.check(L.ddfmm_set_tiles(h, nl, bad))
        with pytest.raises(_lib.DdfmmError, match="entries"):
            _lib.check(L.ddfmm_set_tiles(h, nl - 1, bad))
    finally:
        L.ddfmm_destroy(h)


def test_hbm_budget_sets_the_slab(oplib):
    """Under an HBM budget the context keeps the tables and as many right-hand sides'
    work buffers as fit (the slab); below tables + one right-hand side, setup fails and
    names the minimum."""
    from ddfmm_amd.wave3d import Mlib3d, Wave3d, MIXED

    pos = geometry.sphere(20000, 8.0, 3)
    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": oplib[0], "-mlib3d_hdname": oplib[1]})

    def make(budget_gib):
        w = Wave3d(device=-1)
        w.posptr, w.mlibptr = pos, ml
        w.setup({"-wave3d_ACCU": 2, "-wave3d_K": 8.0, "-wave3d_maxlevel": 12,
                 "-wave3d_hbm_budget": budget_gib}, nrhs_max=5, precision=MIXED)
        return w

    w = make(0)
    m = w.memory()
    st = w.stats()
    w.close()
    assert m["slab"] == 5 and m["fixed_bytes"] > 0 and m["per_rhs_bytes"] > 0
    # every leaf-sorted point: densities, potentials and their staging at least
    assert m["per_rhs_bytes"] >= 4 * 20000 * 8
    G = 1 << 30
    for k, want in ((2.5, 2), (4.999, 4), (7.0, 5)):
        w = make((m["fixed_bytes"] + k * m["per_rhs_bytes"]) / G)
        assert w.memory()["slab"] == want, (k, w.memory())
        w.close()
    with pytest.raises(RuntimeError, match="HBM budget"):
        make((m["fixed_bytes"] + 0.5 * m["per_rhs_bytes"]) / G)
    assert st["nhf_in"] > 0
