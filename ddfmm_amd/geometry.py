"""Seeded synthetic inputs: surface point clouds, complex densities, check targets.

The reference samples WRL meshes in MATLAB (matlab/new_data.m:18-245); the workloads
here are analytic surfaces instead, generated reproducibly from a seed:

* sphere of diameter ``d`` centred at the origin, area-uniform via normalised Gaussian
  vectors, radius shrunk by (1 - 1e-12) so that no point lies on the +K/2 face, which
  the tree builder rejects (wave3d_setup.cpp:126-129);
* ellipsoid with semi-axes (a, b, c) and radial perturbation 1 + eps*sin(3 theta)*cos(5 phi),
  area-uniform by rejection against the surface-area element;
* densities (x + i y)/sqrt(2) with x, y ~ N(0, 1);
* check targets: indices sampled without replacement.
"""
from __future__ import annotations

import numpy as np

SHRINK = 1.0 - 1e-12


def sphere(n: int, diameter: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * (0.5 * diameter * SHRINK)


def _ellipsoid_xyz(theta, phi, axes, eps):
    rad = 1.0 + eps * np.sin(3 * theta) * np.cos(5 * phi)
    st = np.sin(theta)
    return np.stack([axes[0] * rad * st * np.cos(phi),
                     axes[1] * rad * st * np.sin(phi),
                     axes[2] * rad * np.cos(theta)], axis=-1)


def _area_element(theta, phi, axes, eps, h=1e-6):
    xt = (_ellipsoid_xyz(theta + h, phi, axes, eps) - _ellipsoid_xyz(theta - h, phi, axes, eps)) / (2 * h)
    xp = (_ellipsoid_xyz(theta, phi + h, axes, eps) - _ellipsoid_xyz(theta, phi - h, axes, eps)) / (2 * h)
    return np.linalg.norm(np.cross(xt, xp), axis=-1)


def ellipsoid(n: int, axes=(32.0, 8.0, 8.0), eps: float = 0.1, seed: int = 3) -> np.ndarray:
    """Area-uniform samples of the perturbed ellipsoid (rejection on dA/(dtheta dphi))."""
    rng = np.random.default_rng(seed)
    axes = np.asarray(axes, dtype=np.float64)
    tg, pg = np.meshgrid(np.linspace(0, np.pi, 721), np.linspace(0, 2 * np.pi, 1441), indexing="ij")
    amax = 1.05 * _area_element(tg, pg, axes, eps).max()
    out = []
    have = 0
    while have < n:
        m = int((n - have) * 1.6) + 1024
        th = rng.uniform(0, np.pi, m)
        ph = rng.uniform(0, 2 * np.pi, m)
        keep = rng.uniform(0, amax, m) < _area_element(th, ph, axes, eps)
        pts = _ellipsoid_xyz(th[keep], ph[keep], axes, eps)
        out.append(pts)
        have += len(pts)
    return np.concatenate(out)[:n] * SHRINK


def densities(n: int, seed: int, nrhs: int = 1) -> np.ndarray:
    """Complex standard normal densities; shape (n,) for one RHS else (nrhs, n).

    RHS ``r`` uses seed ``seed + r`` so that a multi-RHS batch is the stack of the
    single-RHS vectors."""
    cols = []
    for r in range(nrhs):
        rng = np.random.default_rng(seed + r)
        x = rng.standard_normal((2, n))
        cols.append((x[0] + 1j * x[1]) / np.sqrt(2.0))
    d = np.stack(cols)
    return d[0] if nrhs == 1 else d


def check_targets(n: int, count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(10_000 + seed)
    return np.sort(rng.choice(n, size=min(count, n), replace=False)).astype(np.int64)


# Workloads quoted by the benchmark definition: (name, builder, K, kernel list, nrhs).
WORKLOADS = {
    "sphere8_60k": dict(kind="sphere", n=60_000, diameter=8.0, K=8.0, seed=0, nrhs=1, ptsmax=100),
    "sphere32_1m": dict(kind="sphere", n=1_000_000, diameter=32.0, K=32.0, seed=1, nrhs=1, ptsmax=100),
    "sphere16_250k_x64": dict(kind="sphere", n=250_000, diameter=16.0, K=16.0, seed=2, nrhs=64, ptsmax=100),
    "ellipsoid_4m_x8": dict(kind="ellipsoid", n=4_000_000, K=128.0, seed=3, nrhs=8, ptsmax=400, numC=8),
    # the reference's WRL meshes sampled at NPW points per wavelength (new_data.m)
    "f16_k64": dict(kind="wrl", mesh="F16.wrl", K=64.0, NPW=10.0, seed=0, nrhs=1, ptsmax=100),
    "submarine_k64": dict(kind="wrl", mesh="SubmarineJ.wrl", K=64.0, NPW=10.0, seed=0, nrhs=1, ptsmax=100),
}


def workload_points(name: str) -> np.ndarray:
    w = WORKLOADS[name]
    if w["kind"] == "sphere":
        return sphere(w["n"], w["diameter"], w["seed"])
    if w["kind"] == "wrl":
        import os

        from . import wrl
        pts, tri = wrl.read_wrl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes", w["mesh"]))
        return wrl.sample(pts, tri, w["K"], w["NPW"], seed=w["seed"])
    return ellipsoid(w["n"], seed=w["seed"])
