"""Invariants that pin the oracle's tree, interaction lists and direction algebra to the
mathematics rather than to the C++ planner (which is compared with the oracle byte for
byte in test_cpu_host.py):

* every ordered pair of non-empty leaves (target, source), the self pair included, is
  reached exactly once through the U, V, W, X and high-frequency far lists of the
  target's and source's ancestors (the FMM's partition of G; wave3d_setup.cpp:302-443);
* the lists are symmetric: U and V under exchange, W(t) against X(s), far lists under
  exchange with the key recomputed from the reversed centre difference;
* ``nml2dir`` on diagonal and axis-aligned unit vectors, where the strict ``mval <
  |n(d)|`` comparison breaks ties and ``min(max(t, -B), B - 1)`` clamps the minor
  components (wave3d.cpp:43-71), against values worked out by hand;
* the split rule ``np > ptsmax`` (wave3d_setup.cpp:156-157): a box holding exactly
  ptsmax points stays a leaf, one more point splits it.
"""
import math

import numpy as np
import pytest

from ddfmm_amd import geometry
from ddfmm_amd.wave3d import plan_only
from oracle import wave3d_ref as R


def _live(w):
    return {k: b for k, b in w.boxes.items() if b.tag & R.PTS}


def _subtree_leaves(w, live):
    """leaf indices under every live box (the box itself if it is a leaf)"""
    leaves = [k for k, b in live.items() if b.tag & R.TERMINAL]
    lidx = {k: i for i, k in enumerate(leaves)}
    under = {}
    for k in sorted(live, key=lambda k: -k[0]):          # finest level first
        b = live[k]
        if b.tag & R.TERMINAL:
            under[k] = [lidx[k]]
        else:
            acc = []
            for ind in range(8):
                acc += under.get(R.child_key(k, ind), [])
            under[k] = acc
    return leaves, under


def _adaptive(n=5000, seed=7):
    # ellipsoid in a K = 4 domain: HF levels W = 4, 2, 1 and LF leaves of mixed depth
    pos = geometry.ellipsoid(n, axes=(1.8, 0.5, 0.5), eps=0.1, seed=seed)
    return pos, 4.0


@pytest.fixture(scope="module")
def adaptive_oracle(ops):
    pos, K = _adaptive()
    return R.Wave3dOracle(pos, K, *ops, ptsmax=20).setup()


@pytest.fixture(scope="module")
def sphere_oracle(ops):
    pos = geometry.sphere(8000, 4.0, 11)
    return R.Wave3dOracle(pos, 4.0, *ops, ptsmax=30).setup()


def _cover_counts(w):
    live = _live(w)
    leaves, under = _subtree_leaves(w, live)
    L = len(leaves)
    cnt = np.zeros((L, L), dtype=np.int64)
    kinds = {"U": 0, "V": 0, "W": 0, "X": 0, "far": 0}

    def add(tbox, sbox, kind):
        assert sbox in live, f"{kind}-list source {sbox} has no points"
        ti = np.asarray(under[tbox])
        si = np.asarray(under[sbox])
        cnt[np.ix_(ti, si)] += 1
        kinds[kind] += 1

    for k, b in live.items():
        for s in b.U:
            add(k, s, "U")
        for s in b.V:
            add(k, s, "V")
        for s in b.W:
            add(k, s, "W")
        for s in b.X:
            add(k, s, "X")
        for lst in b.fnde.values():
            for s in lst:
                add(k, s, "far")
    return cnt, kinds


@pytest.mark.parametrize("which", ["adaptive", "sphere"])
def test_lists_cover_every_leaf_pair_once(which, adaptive_oracle, sphere_oracle):
    w = adaptive_oracle if which == "adaptive" else sphere_oracle
    cnt, kinds = _cover_counts(w)
    assert cnt.shape[0] > 50
    assert kinds["U"] > 0 and kinds["V"] > 0 and kinds["far"] > 0
    if which == "adaptive":          # the geometry is chosen to exercise the cross-level lists
        assert kinds["W"] > 0 and kinds["X"] > 0
    missing = int((cnt == 0).sum())
    double = int((cnt > 1).sum())
    assert missing == 0 and double == 0, f"{missing} leaf pairs never interact, {double} more than once"


def test_cover_detects_a_dropped_entry(adaptive_oracle, ops):
    """The cover check fails when one V entry is removed (a planted mistake)."""
    pos, K = _adaptive()
    w = R.Wave3dOracle(pos, K, *ops, ptsmax=20).setup()
    k = next(k for k, b in _live(w).items() if b.V)
    w.boxes[k].V = w.boxes[k].V[1:]
    cnt, _ = _cover_counts(w)
    assert int((cnt == 0).sum()) > 0


def test_lists_are_symmetric(adaptive_oracle):
    w = adaptive_oracle
    live = _live(w)
    for k, b in live.items():
        for s in b.U:
            assert k in live[s].U
        for s in b.V:
            assert k in live[s].V
            d = [k[i] - s[i] for i in (1, 2, 3)]
            assert max(abs(x) for x in d) <= 3 and max(abs(x) for x in d) >= 2
        for s in b.W:
            assert k in live[s].X, "W(t) contains s but X(s) lacks t"
            assert s[0] > k[0]
        for s in b.X:
            assert k in live[s].W
            assert s[0] < k[0]
        cc = w.center(k)
        Wd = w.width(k)
        for d, lst in b.fnde.items():
            for s in lst:
                assert s[0] == k[0]
                diff = cc - w.center(s)
                nrm = math.sqrt(float(diff @ diff))
                assert nrm >= Wd * Wd + Wd - 1e-12
                assert R.nml2dir(diff / nrm, Wd, 4) == d
                back = R.nml2dir(-diff / nrm, Wd, 4)
                assert k in live[s].fnde.get(back, [])


@pytest.mark.parametrize("n,W,expect", [
    # ties between the major axes go to the lowest index (strict mval < |n(d)|); the minor
    # components at 45 degrees sit exactly on a sector boundary, floor gives B and the clamp
    # takes it back to B - 1 (C = 4W, B = C / 2)
    ((1, 1, 0), 1, (4, 3, 1)),
    ((1, 1, 1), 1, (4, 3, 3)),
    ((0, 1, 1), 1, (1, 4, 3)),
    ((1, -1, 0), 1, (4, -3, 1)),
    ((-1, 1, 0), 1, (-4, 3, 1)),
    ((-1, -1, -1), 2, (-8, -7, -7)),
    ((1, 1, 0), 4, (16, 15, 1)),
    # axis-aligned: the minor components are atan(0) = 0 -> sector 0 -> 1 (also for -0.0)
    ((0, 0, 1), 1, (1, 1, 4)),
    ((0, 0, -1), 2, (1, 1, -8)),
    ((-1, 0, 0), 1, (-4, 1, 1)),
    # just inside a sector: atan(tan(pi/8) - 1e-9) < pi/8 -> sector 0 -> 1; just past it -> 3
    ((1, math.tan(math.pi / 8) - 1e-9, 0), 1, (4, 1, 1)),
    ((1, math.tan(math.pi / 8) + 1e-9, 0), 1, (4, 3, 1)),
    ((1, -math.tan(math.pi / 8) + 1e-9, 0), 1, (4, -1, 1)),
])
def test_nml2dir_ties_and_clamps(n, W, expect):
    v = np.asarray(n, dtype=np.float64)
    v = v / np.linalg.norm(v)
    assert R.nml2dir(v, W, 4) == expect


def test_split_boundary_is_strict():
    """K = 1: the root (W = 1) always splits; a W = 1/2 child holding exactly ptsmax
    points is a leaf, with ptsmax + 1 it splits (wave3d_setup.cpp:156-157)."""
    rng = np.random.default_rng(0)
    base = rng.uniform(0.05, 0.45, size=(11, 3))      # all inside the child box [0, 1/2)^3
    for npts, ptsmax, split in ((10, 10, False), (11, 10, True), (11, 11, False)):
        pos = base[:npts]
        w = R.Wave3dOracle(pos, 1.0, None, None, ptsmax=ptsmax).setup()
        child = (1, 1, 1, 1)
        assert w.boxes[child].tag & R.PTS
        assert bool(w.boxes[child].tag & R.TERMINAL) == (not split)
        leaves = [k for k, b in _live(w).items() if b.tag & R.TERMINAL]
        assert (len(leaves) > 1) == split and all(k[0] == (2 if split else 1) for k in leaves)


def test_split_boundary_planner(oplib):
    """The C++ planner applies the same strict rule."""
    rng = np.random.default_rng(0)
    base = rng.uniform(0.05, 0.45, size=(11, 3))
    for npts, ptsmax, leaves in ((10, 10, 1), (11, 11, 1)):
        st, _ = plan_only(base[:npts], 1.0, *oplib, ptsmax=ptsmax)
        assert st["nleaf"] == leaves
    st, _ = plan_only(base, 1.0, *oplib, ptsmax=10)
    assert st["nleaf"] > 1
