"""Driver with the reference's command line and file formats (tt.cpp:43-210).

    python -m ddfmm_amd.tt -posfile pos -denfile den -geomprtn geomprtn -chkfile chk \\
        -valfile val -kernel 0 -mlib3d_ldname helm3d_ld_2.bin -mlib3d_hdname helm3d_hd_2_4.bin \\
        -wave3d_ACCU 2 -wave3d_NPQ 4 -wave3d_K 8 -wave3d_ptsmax 100 -wave3d_maxlevel 12

Files are read under ``data/`` exactly as SeparateRead/SharedRead name them
(parallel.cpp:24-99): per-rank ParVec files ``data/<name>_<rank>_<size>`` (one rank here:
``_0_1``) and shared files ``data/<name>``.  ``-dataroot`` changes ``data``.  Prints the
RESULT block of tt.cpp:189-197 (Ta eval seconds, Td extrapolated direct-sum seconds,
Rt = Td/Ta, Ea = relative error of check()).  Extra options: ``-precision fp64|mixed``.

``python -m ddfmm_amd.tt --make-data DIR --workload sphere8_60k`` writes such a dataset
(plus the operator files) from the seeded generators.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import geometry, mlib_gen, serialize


def parse(argv):
    if len(argv) % 2:
        raise SystemExit("options come in '-key value' pairs")
    return {argv[i]: argv[i + 1] for i in range(0, len(argv), 2)}


def run(opts: dict, out=sys.stdout) -> dict:
    from .wave3d import FP64, MIXED, Kernel3d, Mlib3d, Wave3d

    root = opts.get("-dataroot", "data")

    def need(key):
        if key not in opts:
            raise SystemExit(f"Missing option {key}")
        return opts[key]

    keys, pos, own = serialize.read_parvec_points(os.path.join(root, need("-posfile") + "_0_1"))
    dkeys, den = serialize.read_parvec_cpx(os.path.join(root, need("-denfile") + "_0_1"))
    if not np.array_equal(keys, dkeys):
        raise SystemExit("density keys do not match the point keys")
    order = np.argsort(keys)
    pos, den, keys = pos[order], den[order], keys[order]
    kernel = Kernel3d(int(opts.get("-kernel", 0)))
    ml = Mlib3d("mlib3d_")
    ml.kernel = kernel
    ml.setup({k: os.path.join(root, v) if k in ("-mlib3d_ldname", "-mlib3d_hdname") else v
              for k, v in opts.items()})
    geom = serialize.read_int_tns(os.path.join(root, need("-geomprtn")))
    w = Wave3d("wave3d_")
    w.posptr, w.kernel, w.mlibptr, w.geomprtn = pos, kernel, ml, geom
    prec = MIXED if opts.get("-precision", "fp64") == "mixed" else FP64
    t0 = time.perf_counter()
    w.setup(opts, precision=prec)
    print(f"wave setup used {time.perf_counter() - t0:.3f}secs", file=out)
    t0 = time.perf_counter()
    val = w.eval(den)
    ta = time.perf_counter() - t0
    print(f"wave eval used {ta:.3f}secs", file=out)
    if "-valfile" in opts:
        serialize.write_parvec_cpx(os.path.join(root, opts["-valfile"] + "_0_1"), val)
    res = {"K": w.K, "Ta": ta}
    if "-chkfile" in opts:
        chk = serialize.read_int_vec(os.path.join(root, opts["-chkfile"]))
        idx = np.searchsorted(keys, chk)
        t0 = time.perf_counter()
        ea = w.check(den, val, idx)
        tc = time.perf_counter() - t0
        td = tc * len(keys) / len(chk)
        res.update(Td=td, Rt=td / ta, Ea=ea)
        print("----------------------", file=out)
        print("RESULT", file=out)
        print(f"K  {w.K:.2e}", file=out)
        print(f"Ta {ta:.2e}", file=out)
        print(f"Td {td:.2e}", file=out)
        print(f"Rt {td / ta:.2e}", file=out)
        print(f"Ea {ea:.6e}", file=out)
        print("----------------------", file=out)
    w.close()
    return res


def make_data(root: str, workload: str, nchk: int = 1000, kernel: int = 0):
    """Write pos/den/geomprtn/chk (reference formats) and the operator files under root."""
    os.makedirs(root, exist_ok=True)
    wl = geometry.WORKLOADS[workload]
    pos = geometry.workload_points(workload)
    den = geometry.densities(len(pos), wl["seed"])
    serialize.write_parvec_points(os.path.join(root, "pos_0_1"), pos)
    serialize.write_parvec_cpx(os.path.join(root, "den_0_1"), den)
    numC = wl.get("numC", 1)
    serialize.write_int_tns(os.path.join(root, "geomprtn"), np.zeros((numC,) * 3, dtype=np.int32))
    serialize.write_int_vec(os.path.join(root, "chk"), geometry.check_targets(len(pos), nchk, wl["seed"]))
    ld, hd = mlib_gen.ensure_library(root, ktype=kernel, accu=2, NPQ=4, hf_max_W=4, lf_min_W=2.0 ** -3)
    return pos, den, os.path.basename(ld), os.path.basename(hd)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--make-data":
        a = parse(["-root"] + sys.argv[2:3] + sys.argv[3:])
        make_data(sys.argv[2], a.get("--workload", "sphere8_60k"))
    else:
        run(parse(sys.argv[1:]))
