"""Surface meshes in the reference's WRL subset and point sampling on them.

* ``read_wrl`` follows ReadWrl (file_io.cpp:36-107) and new_readwrl.m: a ``points``
  block of x y z rows, a ``coords`` block of index rows ending in -1 (triangles, or
  quadrilaterals split into the triangles (a, b, c) and (d, a, c)), and the points
  normalised to [-1, 1] by the bounding box (centre at the box midpoint, divided by the
  largest half-extent).
* ``sample`` follows new_data.m:18-196: the mesh scaled by 0.875 K/2, each triangle gets
  ceil(area * NPW^2) points (SamplesInArea, file_io.cpp:205-212) at barycentric
  coordinates drawn from a library of 5000 uniform pairs with sum < 1, seeded here.
* ``write_dataset`` writes pos/den/geomprtn/chk in the reference layout for one rank
  (the cell partition of a single device is all zeros).
"""
from __future__ import annotations

import os

import numpy as np

from . import serialize


def read_wrl(path: str):
    """(points (n, 3) normalised to [-1, 1], triangles (m, 3) zero-based indices)."""
    with open(path) as f:
        tokens = f.read().split()
    if not tokens or tokens[0] != "points":
        raise ValueError(f"{path}: did not find 'points' at the beginning of the geometry file")
    try:
        k = tokens.index("coords")
    except ValueError:
        raise ValueError(f"{path}: did not find 'coords' in the geometry file") from None
    vals = np.array(tokens[1:k], dtype=np.float64)
    if vals.size % 3:
        raise ValueError(f"{path}: point block is not a multiple of 3 values")
    pts = vals.reshape(-1, 3)
    idx = np.array(tokens[k + 1:], dtype=np.int64)
    if idx.size >= 4 and idx[3] == -1:
        per = 3
    elif idx.size >= 5 and idx[4] == -1:
        per = 4
    else:
        raise ValueError(f"{path}: expected -1 after each face")
    if idx.size % (per + 1):
        raise ValueError(f"{path}: face block is not a multiple of {per + 1} values")
    faces = idx.reshape(-1, per + 1)
    if np.any(faces[:, -1] != -1):
        raise ValueError(f"{path}: expected -1 in the last column of every face")
    faces = faces[:, :per]
    if per == 4:   # quadrilateral (a, b, c, d) -> (a, b, c), (d, a, c), interleaved
        tri = np.empty((2 * len(faces), 3), dtype=np.int64)
        tri[0::2] = faces[:, [0, 1, 2]]
        tri[1::2] = faces[:, [3, 0, 2]]
        faces = tri
    if faces.min() < 0 or faces.max() >= len(pts):
        raise ValueError(f"{path}: face index out of range")
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    mid = (lo + hi) / 2
    half = np.max(hi - lo) / 2
    return (pts - mid) / half, faces


def triangle_samples(p1: np.ndarray, p2: np.ndarray, p3: np.ndarray, NPW: float) -> np.ndarray:
    """ceil(area * NPW^2) per triangle, area by Heron's formula (file_io.cpp:205-212)."""
    a = np.linalg.norm(p1 - p2, axis=-1)
    b = np.linalg.norm(p2 - p3, axis=-1)
    c = np.linalg.norm(p3 - p1, axis=-1)
    s = (a + b + c) / 2
    area = np.sqrt(np.maximum(s * (s - a) * (s - b) * (s - c), 0.0))
    return np.ceil(area * NPW * NPW).astype(np.int64)


def sample(points: np.ndarray, tris: np.ndarray, K: float, NPW: float, seed: int = 0) -> np.ndarray:
    """Points on the mesh scaled to 0.875 K/2 (new_data.m:21, 148-171)."""
    rng = np.random.default_rng(seed)
    P = points * (K / 2 * 0.875)
    p1, p2, p3 = P[tris[:, 0]], P[tris[:, 1]], P[tris[:, 2]]
    counts = triangle_samples(p1, p2, p3, NPW)
    lib = rng.random((5000, 2))
    lib = lib[lib.sum(axis=1) < 1]
    lib = np.concatenate([lib, 1 - lib.sum(axis=1, keepdims=True)], axis=1)
    owner = np.repeat(np.arange(len(tris)), counts)
    bary = lib[rng.integers(0, len(lib), size=len(owner))]
    return bary[:, :1] * p1[owner] + bary[:, 1:2] * p2[owner] + bary[:, 2:3] * p3[owner]


def write_dataset(root: str, pos: np.ndarray, K: float, numC: int = 1, nchk: int = 20, seed: int = 0):
    """pos/den/geomprtn/chk files of one rank in the reference layout (new_data.m:200-242)."""
    os.makedirs(root, exist_ok=True)
    rng = np.random.default_rng(seed + 1)
    n = len(pos)
    # densities randn + i*rand, as new_data.m:201 draws them
    den = rng.standard_normal(n) + 1j * rng.random(n)
    serialize.write_parvec_points(os.path.join(root, "pos_0_1"), pos)
    serialize.write_parvec_cpx(os.path.join(root, "den_0_1"), den)
    serialize.write_int_tns(os.path.join(root, "geomprtn"), np.zeros((numC,) * 3, dtype=np.int32))
    chk = np.sort(rng.choice(n, size=min(nchk, n), replace=False)).astype(np.int32)
    serialize.write_int_vec(os.path.join(root, "chk"), chk)
    return den, chk
