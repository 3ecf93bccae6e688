"""Reference file formats used by the tt-style driver: ParVec point/density files,
IntNumTns geomprtn, IntNumVec check keys (serialize.hpp; new_data.m:150-242)."""
import numpy as np

from ddfmm_amd import geometry, serialize


def test_parvec_and_int_formats_roundtrip(tmp_path):
    pos = geometry.sphere(500, 4.0, 1)
    den = geometry.densities(500, 1)
    serialize.write_parvec_points(tmp_path / "pos_0_1", pos)
    serialize.write_parvec_cpx(tmp_path / "den_0_1", den)
    keys, p, own = serialize.read_parvec_points(tmp_path / "pos_0_1")
    assert np.array_equal(keys, np.arange(500)) and np.array_equal(p, pos) and own == [0, 500]
    k2, d = serialize.read_parvec_cpx(tmp_path / "den_0_1")
    assert np.array_equal(k2, keys) and np.array_equal(d, den)
    g = np.arange(8, dtype=np.int32).reshape(2, 2, 2)
    serialize.write_int_tns(tmp_path / "geomprtn", g)
    assert np.array_equal(serialize.read_int_tns(tmp_path / "geomprtn"), g)
    raw = np.fromfile(tmp_path / "geomprtn", dtype=np.int32)
    assert list(raw[:3]) == [2, 2, 2] and raw[3 + 1] == g[1, 0, 0]      # first index fastest
    serialize.write_int_vec(tmp_path / "chk", [5, 7, 11])
    assert list(serialize.read_int_vec(tmp_path / "chk")) == [5, 7, 11]
