"""WRL mesh ingestion (ddfmm_amd/wrl.py) against the reference's own meshes and the
counts its file_io_test.cpp states, plus properties of the sampling that the mathematics
fixes (normalisation, per-triangle counts, points on their triangles)."""
import os

import numpy as np
import pytest

from ddfmm_amd import serialize, wrl

GOLD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ddfmm_amd", "meshes")


@pytest.mark.parametrize("name,npts,ntri", [("sphere.wrl", 1681, 3200), ("F16.wrl", 1165, 2162),
                                            ("SubmarineJ.wrl", 459, 640)])
def test_reference_mesh_counts(name, npts, ntri):
    pts, tri = wrl.read_wrl(os.path.join(GOLD, name))
    assert pts.shape == (npts, 3) and tri.shape == (ntri, 3)
    # normalised to [-1, 1]: the largest extent spans exactly [-1, 1], centred
    ext = pts.max(axis=0) - pts.min(axis=0)
    assert np.isclose(ext.max(), 2.0) and np.all(np.abs(pts) <= 1 + 1e-12)
    assert np.allclose((pts.max(axis=0) + pts.min(axis=0)) / 2, 0, atol=1e-12)


def test_quads_split_into_triangles(tmp_path):
    p = tmp_path / "quad.wrl"
    p.write_text("points\n0 0 0\n2 0 0\n2 2 0\n0 2 0\ncoords\n0 1 2 3 -1\n")
    pts, tri = wrl.read_wrl(str(p))
    assert tri.tolist() == [[0, 1, 2], [3, 0, 2]]
    assert np.allclose(pts[2], [1, 1, 0])
    with pytest.raises(ValueError, match="coords"):
        (tmp_path / "bad.wrl").write_text("points\n0 0 0\n")
        wrl.read_wrl(str(tmp_path / "bad.wrl"))


def test_sampling_counts_and_support():
    pts, tri = wrl.read_wrl(os.path.join(GOLD, "sphere.wrl"))
    K, NPW = 8.0, 10.0
    s = wrl.sample(pts, tri, K, NPW, seed=3)
    P = pts * (K / 2 * 0.875)
    counts = wrl.triangle_samples(P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]], NPW)
    assert len(s) == counts.sum()
    # the unit-sphere mesh scaled to radius 3.5: every sample within the shell the flat
    # triangles span, so inside the K = 8 domain
    r = np.linalg.norm(s, axis=1)
    assert r.max() <= 3.5 * (1 + 1e-12) and r.min() > 3.4 and np.all(np.abs(s) < K / 2)
    # an equilateral triangle of side 1 at NPW = 10 gets ceil(sqrt(3)/4 * 100) = 44 points
    e = wrl.triangle_samples(np.array([[0.0, 0, 0]]), np.array([[1.0, 0, 0]]),
                             np.array([[0.5, np.sqrt(3) / 2, 0]]), 10.0)
    assert e.tolist() == [44]


def test_dataset_files_round_trip(tmp_path):
    pts, tri = wrl.read_wrl(os.path.join(GOLD, "SubmarineJ.wrl"))
    s = wrl.sample(pts, tri, 16.0, 4.0, seed=1)
    den, chk = wrl.write_dataset(str(tmp_path), s, 16.0, numC=2)
    keys, back = serialize.read_parvec_points(str(tmp_path / "pos_0_1"))[:2]
    assert np.array_equal(keys, np.arange(len(s))) and np.array_equal(back, s)
    dk, dd = serialize.read_parvec_cpx(str(tmp_path / "den_0_1"))
    assert np.array_equal(dd, den)
    assert serialize.read_int_tns(str(tmp_path / "geomprtn")).shape == (2, 2, 2)
    assert np.array_equal(serialize.read_int_vec(str(tmp_path / "chk")), chk)
