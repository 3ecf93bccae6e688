"""CPU checks that pin the oracle (and the operator precompute) to mathematics:
closed forms, invariants, brute force, and planted mistakes that must be caught."""
import math

import numpy as np
import pytest

from ddfmm_amd import geometry, mlib_gen, serialize
from oracle import wave3d_ref as R


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- kernel (A1)
def test_kernel_closed_form_values():
    t = np.array([[0.0, 0.0, 0.0]])
    s = np.array([[0.25, 0, 0], [0, 0.5, 0], [0, 0, 1.0], [0, 0, 0], [0, 0, 1e-9]])
    g = R.kernel(t, s, 0)[0]
    assert abs(g[0] - 4j) < 1e-14             # e^{i pi/2} / 0.25
    assert abs(g[1] - (-2.0)) < 1e-14         # e^{i pi} / 0.5
    assert abs(g[2] - 1.0) < 1e-14            # e^{2 pi i} / 1
    assert abs(g[3] - 1.0) < 1e-14            # r^2 < 1e-16 => r := 1 (kernel3d.cpp:47-49)
    assert abs(g[4] - 1.0) < 1e-14
    e = R.kernel(t, s, 1)[0]
    assert abs(e[0] - 1j) < 1e-14 and abs(e[1] + 1) < 1e-14
    with pytest.raises(ValueError):
        R.kernel(t, s, 2)


def test_direct_sum_c_vs_numpy():
    pos = geometry.sphere(700, 4.0, 9)
    den = geometry.densities(700, 9, nrhs=2)
    for kt in (0, 1):
        ref = np.stack([R.kernel(pos[:50], pos, kt) @ den[r] for r in range(2)])
        assert rel(R.direct_sum(pos[:50], pos, den, kt), ref) < 1e-13


# ---------------------------------------------------------------- direction algebra (A16)
@pytest.mark.parametrize("W", [1, 2, 4])
def test_direction_algebra(W):
    NPQ = 4
    C = NPQ * W
    odd = [v for v in range(-C + 1, C, 2)]
    dirs = []
    for ax in range(3):
        for sg in (-1, 1):
            for a in odd:
                for b in odd:
                    d = [a, b]
                    d.insert(ax, sg * C)
                    dirs.append(tuple(d))
    assert len(dirs) == 6 * C * C                       # 6 (NPQ W)^2 directions
    for d in dirs:
        for cd in R.child_dirs(d):
            assert R.parent_dir(cd) == d
        srt, sgn, prm = R.sort_dir(d)
        assert srt[0] <= srt[1] <= srt[2] == C
        # the shuffle maps the canonical direction onto the actual one
        back = R.shuffle(prm, sgn, np.array(srt, dtype=float)[:, None])[0]
        assert tuple(int(x) for x in back) == d
    # nml2dir: the centre of wedge d maps back to d
    for d in dirs:
        v = np.array(d, dtype=float)
        v = np.tan(v / C * math.pi / 4)
        v /= np.linalg.norm(v)
        assert R.nml2dir(v, W, NPQ) == d


# ---------------------------------------------------------------- operators (A6/A7/A12/A19)
def test_lf_pinv_and_far_field(ops):
    ld, _ = ops
    for W in (0.5, 0.25):
        e = ld[W]
        uep, ucp = e["uep"].T, e["ucp"].T
        V, invs, UH = e["uc2ue"]
        G = R.kernel(ucp, uep, 0)
        pinv = V @ (invs * UH)
        assert rel(G @ pinv @ G, G) < 1e-7
        # equivalent density reproduces the far field of random sources inside the box
        rng = np.random.default_rng(1)
        src = rng.uniform(-W / 2, W / 2, (40, 3))
        f = rng.standard_normal(40) + 1j * rng.standard_normal(40)
        eq = pinv @ (R.kernel(ucp, src, 0) @ f)
        d = rng.standard_normal((30, 3))
        far = d / np.linalg.norm(d, axis=1, keepdims=True) * (2.0 * W) + rng.uniform(-0.2, 0.2, (30, 3)) * W
        far = far * 1.6  # well separated targets (distance >= 2W)
        assert rel(R.kernel(far, uep, 0) @ eq, R.kernel(far, src, 0) @ f) < 1e-5


def test_lf_m2l_fft_equals_dense(ops):
    """FFT-diagonal V-list translation == dense G(dcp + offset*W, uep) (wave3d_eval.cpp:602-661)."""
    ld, _ = ops
    W, P = 0.25, 6
    e = ld[W]
    uep = e["uep"].T
    step = W / (P - 1)
    gi = np.rint((uep + W / 2) / step).astype(int) + P
    rng = np.random.default_rng(2)
    x = rng.standard_normal(len(uep)) + 1j * rng.standard_normal(len(uep))
    for off in [(2, 0, 0), (-3, 1, 2), (0, -2, -3), (3, 3, 3)]:
        g = np.zeros((12, 12, 12), complex)
        g[gi[:, 0], gi[:, 1], gi[:, 2]] = x
        y = np.fft.ifftn(np.fft.fftn(g) * e["ue2dc"][off])[gi[:, 0], gi[:, 1], gi[:, 2]]
        dense = R.kernel(uep + np.array(off) * W, uep, 0) @ x
        assert rel(y, dense) < 1e-12


@pytest.mark.parametrize("W", [1, 2])
def test_hf_directional_reproduction(ops, W):
    """Directional equivalent densities reproduce the far field inside the wedge."""
    _, hd = ops
    rng = np.random.default_rng(3)
    for n, e in hd[float(W)].items():
        uep, ucp = e["uep"].T, e["ucp"].T
        E1, E2, E3 = e["uc2ue"]
        src = rng.uniform(-W / 2, W / 2, (60, 3))
        f = rng.standard_normal(60) + 1j * rng.standard_normal(60)
        eq = E1 @ (E2 @ (E3 @ (R.kernel(ucp, src, 0) @ f)))
        C = 4 * W
        D = W * W + W
        ang = (np.array(n[:2], float) + rng.uniform(-1, 1, (40, 2))) / C * math.pi / 4
        v = np.stack([np.tan(ang[:, 0]), np.tan(ang[:, 1]), np.ones(40)], 1)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        trg = v * rng.uniform(D + W, 4 * D, (40, 1))
        assert rel(R.kernel(trg, uep, 0) @ eq, R.kernel(trg, src, 0) @ f) < 2e-5


def test_serialize_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    e = dict(uep=rng.standard_normal((3, 5)), ucp=rng.standard_normal((3, 4)),
             uc2ue=[rng.standard_normal((5, 4)) + 1j, rng.standard_normal((4, 1)) + 0j, rng.standard_normal((4, 4)) + 2j],
             ue2dc={(2, 0, -3): rng.standard_normal((12, 12, 12)) + 1j})
    serialize.write_ld(tmp_path / "ld.bin", {0.5: e})
    back = serialize.read_ld(tmp_path / "ld.bin")[0.5]
    assert np.array_equal(back["uep"], e["uep"]) and np.array_equal(back["ucp"], e["ucp"])
    assert all(np.array_equal(a, b) for a, b in zip(back["uc2ue"], e["uc2ue"]))
    assert list(back["ue2dc"]) == [(2, 0, -3)] and np.array_equal(back["ue2dc"][(2, 0, -3)], e["ue2dc"][(2, 0, -3)])
    h = {1.0: {(1, 3, 4): dict(uep=e["uep"], ucp=e["ucp"], uc2ue=e["uc2ue"])}}
    serialize.write_hd(tmp_path / "hd.bin", h)
    hb = serialize.read_hd(tmp_path / "hd.bin")
    assert list(hb[1.0]) == [(1, 3, 4)] and np.array_equal(hb[1.0][(1, 3, 4)]["uc2ue"][2], e["uc2ue"][2])


# ---------------------------------------------------------------- end to end (A4-A15, A20)
def test_oracle_fmm_vs_direct(ops):
    K, N = 4.0, 5000
    pos = geometry.sphere(N, K, 11)
    den = geometry.densities(N, 11)
    val = R.Wave3dOracle(pos, K, *ops).setup().eval(den)
    chk = geometry.check_targets(N, 300, 11)
    assert rel(val[chk], R.direct_sum(pos[chk], pos, den)) < 1e-5


def test_planted_mistakes_are_caught(ops, monkeypatch):
    """A dropped term or a sign error in the oracle pushes it far from the direct sum."""
    K, N = 4.0, 3000
    pos = geometry.sphere(N, K, 12)
    den = geometry.densities(N, 12)
    chk = geometry.check_targets(N, 200, 12)
    exact = R.direct_sum(pos[chk], pos, den)
    # wrong child shift sign in M2M/L2L
    orig = R.shifted_point
    monkeypatch.setattr(R, "shifted_point", lambda ind, W: -orig(ind, W))
    bad = R.Wave3dOracle(pos, K, *ops).setup().eval(den)
    assert rel(bad[chk], exact) > 1e-3
    monkeypatch.setattr(R, "shifted_point", orig)
    # HF M2L direction reversed (source - target)
    orig_dir = R.nml2dir
    w = R.Wave3dOracle(pos, K, *ops).setup()
    for b in w.boxes.values():
        b.fnde = {tuple(-x for x in d): v for d, v in b.fnde.items()}
    bad2 = w.eval(den)
    assert rel(bad2[chk], exact) > 1e-3
    assert orig_dir is R.nml2dir
