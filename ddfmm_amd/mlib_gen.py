"""Setup-time precompute of the translation operators (FP64), written in the reference's
operator-library format so that ``Mlib3d`` reads them unchanged.

This restates the offline MATLAB generators:

* low frequency, per box width W < 1 (matlab/aug3d_lowdata.m:18-117):
  equivalent surface ``uep`` = boundary nodes of a P^3 grid over [-W/2, W/2]^3,
  check surface ``ucp`` = the same over [-3W/2, 3W/2]^3, the check->equivalent
  pseudo-inverse of G(ucp, uep) as an SVD triple {V, 1/s, U^H} truncated at
  1e-16*s_1, and the 316 FFT'd M2L tensors for centre offsets in [-3,3]^3 \\ [-1,1]^3
  sampled on the (2P)^3 grid [0..P-1, -P..-1]*W/(P-1) + offset*W;
* high frequency, per width W >= 1 and canonical direction (a, b, C), a <= b < C odd,
  C = NPQ*W (matlab/aug3d_hghdata.m:18-286): a randomized pivoted-QR skeletonisation of
  G(wedge targets, child equivalent points) giving the directional equivalent points
  ``uep`` (source skeleton), check points ``ucp`` (target skeleton) and the chain
  {E1 = R1^-1, E2 = G, E3 = R2^-H}.  Level W uses level W/2's directional ``uep`` as
  child points (aug3d_hghdata.m:62-91, predecessor by aug3d_hghdata_predir.m).

The MATLAB sampling is unseeded; here every direction draws from its own generator
seeded by (seed, W, direction), so a library is reproducible.  Both kernels
e^{ikr}/r (type 0) and e^{ikr} (type 1) are supported with k = 2*pi (kernel3d.m:21-32,
kernel3d.cpp:26-107).
"""
from __future__ import annotations

import os
import zlib
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import scipy.linalg as sla
from scipy.spatial import ConvexHull
from threadpoolctl import threadpool_limits

from . import serialize

TPK = 2.0 * np.pi
KERNEL_HELM = 0
KERNEL_EXPR = 1
KIND_NAME = {KERNEL_HELM: "helm", KERNEL_EXPR: "expr"}


def P_of(accu: int) -> int:
    return {1: 4, 2: 6, 3: 8}[accu]


def gkernel(t: np.ndarray, s: np.ndarray, ktype: int) -> np.ndarray:
    """Dense G(t_i, s_j) for row-point arrays t (m,3), s (n,3); no coincidence guard
    (the operator point sets never coincide, as in kernel3d.m)."""
    d = t[:, None, :] - s[None, :, :]
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    ph = np.exp(1j * TPK * r)
    return ph / r if ktype == KERNEL_HELM else ph


def _ndgrid_pts(g):
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel("F"), Y.ravel("F"), Z.ravel("F")], axis=1)


def cube_surface(L: float, P: int) -> np.ndarray:
    """Boundary nodes of the P^3 grid on [-L/2, L/2]^3, first coordinate fastest."""
    pp = _ndgrid_pts(np.linspace(-L / 2, L / 2, P))
    good = np.max(np.abs(pp), axis=1) == L / 2
    return pp[good]


# --------------------------------------------------------------------------- LF
def lowdata(accu: int, W: float, ktype: int) -> dict:
    P = P_of(accu)
    uep = cube_surface(W, P)
    ucp = cube_surface(3 * W, P)
    Mce = gkernel(ucp, uep, ktype)
    u, s, vh = np.linalg.svd(Mce, full_matrices=False)
    gd = s > 1e-16 * s[0]
    V = vh.conj().T[:, gd]
    invs = (1.0 / s[gd]).astype(np.complex128)[:, None]
    UH = u[:, gd].conj().T
    step = W / (P - 1)
    g = np.concatenate([np.arange(P), np.arange(-P, 0)]) * step
    ue2dc = {}
    for c in range(-3, 4):
        for b in range(-3, 4):
            for a in range(-3, 4):
                if max(abs(a), abs(b), abs(c)) <= 1:
                    continue
                X, Y, Z = np.meshgrid(g + a * W, g + b * W, g + c * W, indexing="ij")
                r = np.sqrt(X * X + Y * Y + Z * Z)
                ph = np.exp(1j * TPK * r)
                Moc = ph / r if ktype == KERNEL_HELM else ph
                ue2dc[(a, b, c)] = np.fft.fftn(Moc)
    return dict(uep=uep.T.copy(), ucp=ucp.T.copy(), uc2ue=[V, invs, UH], ue2dc=ue2dc)


# --------------------------------------------------------------------------- HF
def canonical_dirs(W: int, NPQ: int):
    C = NPQ * W
    odd = range(1, C, 2)
    return sorted({tuple(sorted((a, b, C))) for a in odd for b in odd})


def predir_canonical(n):
    C = max(abs(v) for v in n)
    r = [2 * (((v + C - 1) // 2) // 2) + 1 - C // 2 for v in n]
    r[2] = n[2] // 2
    return tuple(r)


def _hull_idx(p: np.ndarray) -> np.ndarray:
    return np.unique(ConvexHull(p).simplices.ravel())


def _unique_rows(p: np.ndarray) -> np.ndarray:
    return np.unique(p, axis=0)


def _mround(x: float) -> int:
    return int(np.floor(x + 0.5))


def _rand_subset(rng, k: int, n: int) -> np.ndarray:
    """Sorted distinct random indices: sort(ceil(rand(k)*(n-k))) + [1..k], 0-based
    (aug3d_hghdata.m:183-186)."""
    return np.sort(np.ceil(rng.random(k) * (n - k)).astype(np.int64)) + np.arange(k)


_HF_PARAMS = {1: (1e-4, 70, 4, 2400), 2: (1e-6, 120, 5, 2400), 3: (1e-8, 160, 6, 4800)}


def hghdata_dir(accu, NPQ, W, n, pspre, ktype, seed):
    """One canonical direction n = (a, b, C) of width W; pspre = child equivalent
    points (rows).  Returns dict(uep (3,r1), ucp (3,r2), uc2ue=[E1, E2, E3])."""
    EPS, rr, RT, NT = _HF_PARAMS[accu]
    key = f"{seed}:{W}:{n}".encode()
    rng = np.random.default_rng(zlib.crc32(key) + 7919 * seed)
    D = W * W + W
    DP = (2 * W) ** 2 + 2 * W
    C = NPQ * W
    n = np.asarray(n, dtype=np.float64)
    # candidate far-field centres in the first octant and their direction coordinates
    tmp = np.arange(0, 5 * D // W + 1) * float(W)
    sfts = _ndgrid_pts(tmp)
    sr = np.linalg.norm(sfts, axis=1)
    sfts = sfts[(sr >= D - 1e-8) & (sr <= 5 * D)]
    ns = sfts / np.linalg.norm(sfts, axis=1, keepdims=True)
    ns = ns / np.max(np.abs(ns), axis=1, keepdims=True)
    ns = np.arctan(ns) / (np.pi / 4) * C

    ptpre = -pspre
    sgs = _ndgrid_pts(np.array([-W / 4, W / 4]))
    pscol = _unique_rows(np.concatenate([pspre + s for s in sgs]))
    ptcol = _unique_rows(np.concatenate([ptpre + s for s in sgs]))
    ps = pscol
    salw = _hull_idx(ps)

    tol = 1.0 + 1e-8
    good = (np.abs(ns[:, 0] - n[0]) <= tol) & (np.abs(ns[:, 1] - n[1]) <= tol) & \
           (np.abs(ns[:, 2] - n[2]) <= 1e-8)
    sss = sfts[good]
    sgs2 = _ndgrid_pts(np.array([-W / 2, W / 2]))
    play = np.concatenate([(1.0 if np.linalg.norm(c) < DP - 1e-8 else 3.0) * sgs2 + c for c in sss])
    pt0 = play[_hull_idx(play)]
    dis = np.linalg.norm(sss, axis=1)
    mindis = dis.min()
    pt1 = _unique_rows(np.concatenate([ptcol + sss[k] for k in np.nonzero(dis < mindis + 1e-8)[0]]))
    thes = np.arctan2(play[:, 0], play[:, 2])
    phis = np.arctan2(play[:, 1], play[:, 2])
    the_rng = (thes.min(), thes.max())
    phi_rng = (phis.min(), phis.max())
    a, b = n[0], n[1]
    if (the_rng[0] > (a - 1) / C * np.pi / 4 or the_rng[1] < (a + 1) / C * np.pi / 4 or
            phi_rng[0] > (b - 1) / C * np.pi / 4 or phi_rng[1] < (b + 1) / C * np.pi / 4):
        raise RuntimeError(f"wedge sampling does not cover direction {n} at W={W}")
    rho_rng = (mindis - W / 2, 6 * D)
    the = rng.random(NT) * (the_rng[1] - the_rng[0]) + the_rng[0]
    phi = rng.random(NT) * (phi_rng[1] - phi_rng[0]) + phi_rng[0]
    rho = rng.random(NT) * (rho_rng[1] - rho_rng[0]) + rho_rng[0]
    tq = np.sqrt(np.tan(the) ** 2 + np.tan(phi) ** 2 + 1)
    pt2 = np.stack([np.tan(the) / tq * rho, np.tan(phi) / tq * rho, rho / tq], axis=1)
    talw2 = pt2[_hull_idx(pt2)]
    pt = _unique_rows(np.concatenate([pt0, pt1, pt2]))
    talw_pts = _unique_rows(np.concatenate([pt0, pt1, talw2]))
    # indices of talw_pts inside pt (both lexicographically sorted and talw_pts is a subset)
    talw = np.searchsorted(pt.view([("", pt.dtype)] * 3).ravel(),
                           talw_pts.view([("", pt.dtype)] * 3).ravel())

    M, N = len(pt), len(ps)
    sz1, sz2 = rr * RT, rr * RT * RT
    nr = _mround(sz1 * np.sqrt(M / N))
    rs = np.arange(M) if nr > M else _rand_subset(rng, nr, M)
    rs = np.unique(np.concatenate([rs, talw]))
    M1 = gkernel(pt[rs], ps, ktype)
    _, R1, E1 = sla.qr(M1, mode="economic", pivoting=True)
    d1 = np.abs(np.diag(R1))
    idx1 = E1[:len(d1)][d1 > EPS * d1[0]]
    psidx = ps[idx1]
    Q1, R1 = sla.qr(gkernel(pt, psidx, ktype), mode="economic")

    nc = sz1
    cs = np.arange(N) if nc > N else _rand_subset(rng, nc, N)
    cs = np.unique(np.concatenate([cs, salw, idx1]))
    M2 = gkernel(pt, ps[cs], ktype)
    _, R2, E2 = sla.qr(M2.conj().T, mode="economic", pivoting=True)
    d2 = np.abs(np.diag(R2))
    idx2 = E2[:len(d2)][d2 > EPS * d2[0]]
    ptidx = pt[idx2]
    Q2, R2 = sla.qr(gkernel(ptidx, ps, ktype).conj().T, mode="economic")

    nc = sz2
    cs = np.arange(N) if nc > N else _rand_subset(rng, nc, N)
    cs = np.unique(np.concatenate([idx1, cs]))
    nr = sz2
    rs = np.arange(M) if nr > M else _rand_subset(rng, nr, M)
    rs = np.unique(np.concatenate([idx2, rs]))
    M3 = gkernel(pt[rs], ps[cs], ktype)
    G = np.linalg.pinv(Q1[rs]) @ M3 @ np.linalg.pinv(Q2[cs].conj().T)
    E1m = sla.inv(R1)
    E3m = sla.inv(R2.conj().T)
    return dict(uep=psidx.T.copy(), ucp=ptidx.T.copy(), uc2ue=[E1m, G, E3m])


def _hf_job(args):
    accu, NPQ, W, n, pspre, ktype, seed, nthreads = args
    with threadpool_limits(limits=nthreads):
        return n, hghdata_dir(accu, NPQ, W, n, pspre, ktype, seed)


def hghdata(accu, NPQ, W, respre, ktype, seed, workers=1):
    P = P_of(accu)
    jobs = []
    for n in canonical_dirs(W, NPQ):
        if W == 1:
            pspre = cube_surface(W / 2, P)
        else:
            pspre = respre[predir_canonical(n)]["uep"].T
        jobs.append((accu, NPQ, W, n, pspre, ktype, seed, 1 if workers > 1 else None))
    if workers > 1 and len(jobs) > 1:
        with ProcessPoolExecutor(max_workers=min(workers, len(jobs))) as ex:
            out = dict(ex.map(_hf_job, jobs))
    else:
        out = dict(map(_hf_job, jobs))
    return out


# --------------------------------------------------------------------------- files
def library_paths(root: str, ktype: int, accu: int, NPQ: int):
    k = KIND_NAME[ktype]
    return (os.path.join(root, f"{k}3d_ld_{accu}.bin"),
            os.path.join(root, f"{k}3d_hd_{accu}_{NPQ}.bin"))


def ensure_library(root: str, ktype: int = 0, accu: int = 2, NPQ: int = 4,
                   hf_max_W: int = 4, lf_min_W: float = 2.0 ** -3, seed: int = 0,
                   workers: int | None = None, verbose: bool = False):
    """Generate (or extend) the LF/HF operator files under ``root``; returns paths.

    Existing files are kept and only missing widths are added (HF widths are chained,
    so an extension starts from the largest stored width)."""
    os.makedirs(root, exist_ok=True)
    ldp, hdp = library_paths(root, ktype, accu, NPQ)
    workers = workers or min(8, os.cpu_count() or 1)
    ld = serialize.read_ld(ldp) if os.path.exists(ldp) else {}
    want_lf = []
    W = 0.5
    while W >= lf_min_W * (1 - 1e-12):
        want_lf.append(W)
        W /= 2
    missing = [w for w in want_lf if w not in ld]
    if missing:
        for w in missing:
            if verbose:
                print(f"[mlib] LF W={w}", flush=True)
            ld[w] = lowdata(accu, w, ktype)
        serialize.write_ld(ldp, ld)
    hd = serialize.read_hd(hdp) if os.path.exists(hdp) else {}
    changed = False
    W = 1
    prev = None
    while W <= hf_max_W:
        if float(W) not in hd:
            if verbose:
                print(f"[mlib] HF W={W} ({len(canonical_dirs(W, NPQ))} dirs)", flush=True)
            hd[float(W)] = hghdata(accu, NPQ, W, prev, ktype, seed, workers)
            changed = True
        prev = hd[float(W)]
        W *= 2
    if changed:
        serialize.write_hd(hdp, hd)
    return ldp, hdp
