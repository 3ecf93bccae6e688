"""Stage-by-stage GPU parity: intermediate device buffers vs the oracle's intermediates
(LF upward equivalent densities, HF directional equivalent densities, the width-1 HF
check values, LF downward equivalent densities, leaf potentials).  The HF check values
of a level occupy one level-sized region that the next finer level reuses, so after an
evaluation the region holds width 1's, which depend on every coarser level through the
L2L chain.

Equivalent densities come out of ill-conditioned solves (cond G(ucp, uep) ~ 3e8 for the
LF surfaces, cond R1 ~ 1e6 for the HF chain), so two FP64 evaluations that differ only in
summation order agree on them to ~cond * eps: tolerance 1e-7.  Check values and
potentials are well conditioned: tolerance 1e-12."""
import numpy as np
import pytest

from ddfmm_amd import geometry
from ddfmm_amd.wave3d import Kernel3d, Mlib3d, Wave3d
from oracle import wave3d_ref as R

pytestmark = pytest.mark.gpu
BUF_VAL, BUF_LF_UPEQ, BUF_LF_DNEQ, BUF_HF_UP, BUF_HF_DN = 1, 2, 5, 6, 7


def slot_maps(ora, hd, NL):
    live = {k: b for k, b in ora.boxes.items() if b.tag & R.PTS}
    lf = [k for k in live if ora.width(k) < 1 - 1e-12]
    lfslot = {k: i * NL for i, k in enumerate(lf)}
    hdmap = {}
    for k in live:
        if ora.width(k) >= 1 - 1e-12:
            for d in sorted(live[k].outdir):
                hdmap.setdefault(d, ([], []))[0].append(k)
            for d in sorted(live[k].incdir):
                hdmap.setdefault(d, ([], []))[1].append(k)
    byw = {}
    for d in sorted(hdmap):
        byw.setdefault(max(abs(v) for v in d) // 4, []).append(d)
    up, dn = {}, {}
    cu = 0
    for W in sorted(byw):
        cd = 0                      # in-slots are level-local
        for d in byw[W]:
            r1 = hd[float(W)][R.sort_dir(d)[0]]["uep"].shape[1]
            for k in hdmap[d][0]:
                up[(k, d)] = (cu, r1)
                cu += r1
            for k in hdmap[d][1]:
                if W == 1:
                    dn[(k, d)] = (cd, r1)
                cd += r1
    return lfslot, up, dn


def stack_rel(pairs):
    a = np.concatenate([p[0] for p in pairs])
    b = np.concatenate([p[1] for p in pairs])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("K,N", [(4.0, 6000), (8.0, 30000)])
def test_stage_buffers_match_oracle(oplib, ops, K, N):
    pos = geometry.sphere(N, K, 1)
    den = geometry.densities(N, 1)
    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": oplib[0], "-mlib3d_hdname": oplib[1]})
    w = Wave3d()
    w.posptr, w.mlibptr, w.kernel = pos, ml, Kernel3d(0)
    w.setup({"-wave3d_ACCU": 2, "-wave3d_K": K, "-wave3d_maxlevel": 12})
    val = w.eval(den)
    ora = R.Wave3dOracle(pos, K, *ops).setup()
    ref = ora.eval(den)
    NL = 152
    lfslot, up, dn = slot_maps(ora, ops[1], NL)
    res = {}
    res["lf_upeq"] = stack_rel([(w.read_buffer(BUF_LF_UPEQ, s, NL), ora.upeq[k]) for k, s in lfslot.items()])
    res["hf_up"] = stack_rel([(w.read_buffer(BUF_HF_UP, o, r), ora.dirup[kd]) for kd, (o, r) in up.items()])
    res["hf_dn"] = stack_rel([(w.read_buffer(BUF_HF_DN, o, r), ora.dirdn[kd]) for kd, (o, r) in dn.items()])
    res["lf_dneq"] = stack_rel([(w.read_buffer(BUF_LF_DNEQ, s, NL), ora.dneq[k]) for k, s in lfslot.items()])
    res["val"] = float(np.linalg.norm(val - ref) / np.linalg.norm(ref))
    print(f"K={K} N={N} " + " ".join(f"{k}={v:.2e}" for k, v in res.items()), flush=True)
    w.close()
    tol = {"lf_upeq": 1e-7, "hf_up": 1e-7, "lf_dneq": 1e-7, "hf_dn": 1e-12, "val": 1e-12}
    for k, v in res.items():
        assert v < tol[k], (k, v)
