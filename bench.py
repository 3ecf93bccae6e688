#!/usr/bin/env python3
"""Throughput of the directional-FMM evaluation pass on one MI355X.

One step = one Wave3d::eval over the workload (densities and potentials resident in
HBM, eval_device).  Default workload: sphere of diameter 32 wavelengths, N = 1,000,000
area-uniform points (seed 1), one complex-normal density vector, kernel exp(ikr)/r,
ACCU=2 (P=6), NPQ=4, ptsmax=100.  Metric: M point-evaluations/s per RHS, reported
with the relative L2 error at 1000 random targets against an FP64 direct sum.

Prints one JSON line.  Also reports the roofline of the dominant kernel (HIP events on
the evaluation stream) and a CPU baseline (the oracle's numpy FMM on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}
ST_ZERO, ST_GATHER, ST_SCATTER, ST_KAPPLY, ST_GEMM, ST_FFTFWD, ST_VLIST, ST_REDUCE, ST_FORK, ST_JOIN = range(10)


def kernel_class(kind: int, prec: int) -> str:
    if kind == ST_GEMM:
        return "gemm_solve" if prec == 1 else "gemm_translation"
    return {ST_KAPPLY: "kapply", ST_VLIST: "vlist", ST_FFTFWD: "fft_fwd", ST_ZERO: "zero",
            ST_GATHER: "gather", ST_SCATTER: "scatter", ST_REDUCE: "reduce", ST_FORK: "sync", ST_JOIN: "sync"}[kind]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ops_paths, kernel):
    """The oracle (numpy FP64 FMM) on the 60k-point sphere of diameter 8 (K = 8)."""
    from threadpoolctl import threadpool_limits

    from ddfmm_amd import geometry, serialize
    from oracle import wave3d_ref

    cores = min(16, len(os.sched_getaffinity(0)))
    N, K = 60_000, 8.0
    pos = geometry.sphere(N, K, 0)
    den = geometry.densities(N, 0)
    ld, hd = serialize.read_ld(ops_paths[0]), serialize.read_hd(ops_paths[1])
    with threadpool_limits(limits=cores):
        w = wave3d_ref.Wave3dOracle(pos, K, ld, hd, ktype=kernel).setup()
        t0 = time.perf_counter()
        w.eval(den)
        dt = time.perf_counter() - t0
    return {"value": N / dt / 1e6, "unit": "M point-evals/s·RHS", "cores": cores, "kind": "port",
            "sample": f"oracle numpy FMM eval (FP64) on the sphere d=8 N=60000 K=8 workload, "
                      f"one RHS, {dt:.1f} s"}


def serve_bench(names, qps, count, slot, precision, seed=4):
    """Poisson request stream over resident geometries (continuous batching)."""
    from ddfmm_amd import geometry, mlib_gen
    from ddfmm_amd.serve import GpuEngine, Server, latency_stats, poisson_requests
    from ddfmm_amd.wave3d import FP64, MIXED

    ops = mlib_gen.ensure_library(os.path.join(ROOT, "ddfmm_amd", "_ops"), ktype=0, accu=2, NPQ=4,
                                  hf_max_W=4, lf_min_W=2.0 ** -3)
    pts = {n: geometry.workload_points(n) for n in names}

    def make(key):
        wl = geometry.WORKLOADS[key]
        return GpuEngine(key, pts[key], {"-wave3d_ACCU": 2, "-wave3d_K": wl["K"], "-wave3d_ptsmax": wl["ptsmax"],
                                         "-wave3d_maxlevel": 14}, ops, slot,
                         MIXED if precision == "mixed" else FP64, numC=wl.get("numC", 1))

    srv = Server(make, lambda key: 0, hbm_budget=200 << 30)
    t0 = time.perf_counter()
    for n in names:                     # resident before the stream starts (warm cache)
        srv.engines[n] = make(n)
    setup_s = time.perf_counter() - t0
    warm = poisson_requests({n: len(pts[n]) for n in names}, qps=qps, count=len(names), seed=seed + 1)
    srv.run(warm)
    reqs = poisson_requests({n: len(pts[n]) for n in names}, qps=qps, count=count, seed=seed)
    t0 = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t0
    st = latency_stats(reqs)
    evals = sum(r.nrhs * len(pts[r.geometry]) for r in reqs if r.error is None)
    out = {"geometries": names, "qps": qps, "requests": count, "rhs_per_request": "1-16", "slot_rhs": slot,
           "batches": srv.batches, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st.get("mean_ms"),
           "throughput_M_point_evals_per_s": evals / wall / 1e6, "wall_s": wall, "setup_s": setup_s,
           "errors": sum(r.error is not None for r in reqs),
           "resident_GiB": srv.resident_bytes() / 2**30}
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sphere32_1m")
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--nrhs", type=int, default=None)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed", "fp64_fused"])
    ap.add_argument("--calib-copy", action="store_true",
                    help="one 256 MiB device-to-device copy (blit kernel) to calibrate PMC byte counters")
    ap.add_argument("--chk", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve", default="sphere8_60k,sphere16_250k_x64,sphere32_1m,ellipsoid_4m_x8",
                    help="comma-separated geometries for the serving leg ('' to skip)")
    ap.add_argument("--serve-qps", type=float, default=10.0)
    ap.add_argument("--serve-requests", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(one device share per process, no collectives).", file=sys.stderr)
        sys.exit(2)

    from ddfmm_amd import geometry, mlib_gen
    from ddfmm_amd._lib import Hip
    from ddfmm_amd.wave3d import FP64, FP64_FUSED, MIXED, Kernel3d, Mlib3d, Wave3d, direct_sum

    wl = dict(geometry.WORKLOADS[args.workload])
    nrhs = args.nrhs or wl["nrhs"]
    ops = mlib_gen.ensure_library(os.path.join(ROOT, "ddfmm_amd", "_ops"), ktype=args.kernel, accu=2, NPQ=4,
                                  hf_max_W=4, lf_min_W=2.0 ** -3)
    t0 = time.perf_counter()
    pos = geometry.workload_points(args.workload)
    den = geometry.densities(len(pos), wl["seed"], nrhs=nrhs)
    N = len(pos)
    log(f"workload {args.workload}: N={N} nrhs={nrhs} kernel={args.kernel} ({time.perf_counter() - t0:.1f}s)")

    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": ops[0], "-mlib3d_hdname": ops[1]})
    w = Wave3d()
    w.posptr, w.mlibptr, w.kernel = pos, ml, Kernel3d(args.kernel)
    w.geomprtn = np.zeros((wl.get("numC", 1),) * 3, dtype=np.int32)
    t0 = time.perf_counter()
    w.setup({"-wave3d_ACCU": 2, "-wave3d_NPQ": 4, "-wave3d_K": wl["K"], "-wave3d_ptsmax": wl["ptsmax"],
             "-wave3d_maxlevel": 14}, nrhs_max=nrhs,
            precision={"mixed": MIXED, "fp64": FP64, "fp64_fused": FP64_FUSED}[args.precision])
    st = w.stats()
    log(f"setup {time.perf_counter() - t0:.1f}s: leaves={st['nleaf']} m2l_pairs={st['m2l_pairs']} "
        f"v_pairs={st['v_pairs']} device={st['device_bytes'] / 2**30:.2f} GiB")

    hip = Hip()
    if args.calib_copy:
        nb = 256 << 20
        a, b = hip.malloc(nb), hip.malloc(nb)
        hip.copy_d2d(b, a, nb)
        hip.sync()
        hip.free(a)
        hip.free(b)
    den2 = np.ascontiguousarray(np.atleast_2d(den))
    dd = hip.to_device(den2)
    vd = hip.malloc(den2.nbytes)
    for _ in range(args.warmup):
        w.eval_device(dd, vd, nrhs)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.eval_device(dd, vd, nrhs)
    hip.sync()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"{args.steps} steps: {dt * 1e3:.3f} ms/step")

    # accuracy at random targets vs the FP64 direct sum
    val = hip.to_host(vd, np.empty_like(den2))
    chk = geometry.check_targets(N, args.chk, wl["seed"])
    exact = direct_sum(pos[chk], pos, den2, args.kernel)
    err = float(np.linalg.norm(val[:, chk] - exact) / np.linalg.norm(exact))
    log(f"rel.err vs FP64 direct at {len(chk)} targets: {err:.3e}")

    # per-kernel time from HIP events on the evaluation stream
    w.set_timing(True)
    per = {}
    per_stage = {}
    reps = 3
    for _ in range(reps):
        w.eval_device(dd, vd, nrhs)
        for name, ms, work, kind, prec in w.stage_times():
            if kind in (ST_FORK, ST_JOIN):
                continue
            ps = per_stage.setdefault(name, [0.0, 0.0])
            ps[0] += ms / reps
            ps[1] += work * nrhs / reps
            k = kernel_class(kind, prec)
            e = per.setdefault(k, {"ms": 0.0, "launches": 0, "work": 0.0})
            e["ms"] += ms / reps
            e["launches"] += 1
            e["work"] += work * nrhs / reps
    for k in per:
        per[k]["launches"] //= reps
    w.set_timing(False)
    log("top stages (ms, work/ms): " + ", ".join(
        f"{n}={v[0]:.3f}({v[1] / max(v[0], 1e-9) / 1e9:.2f}G/ms)" for n, v in
        sorted(per_stage.items(), key=lambda kv: -kv[1][0])[:12]))
    dom = max(per, key=lambda k: per[k]["ms"])
    t_dom = per[dom]["ms"] * 1e-3
    launches = per[dom]["launches"]
    log("per kernel class (ms/eval, launches): " + ", ".join(
        f"{k}={v['ms']:.3f}/{v['launches']}" for k, v in sorted(per.items(), key=lambda kv: -kv[1]["ms"])))
    mixed = args.precision == "mixed"
    esz = 8 if mixed else 16
    vt = "float2" if mixed else "double2"
    kname = {"gemm_translation": "gemm_kernel<float2,float2,float>" if mixed else "gemm_kernel<double2,double2,double>",
             "gemm_solve": "gemm_kernel<double2,float2,double>" if mixed else "gemm_kernel<double2,double2,double>",
             "kapply": f"kapply_kernel<{vt},{1 if nrhs == 1 else 4}>", "vlist": f"vlist_acc_kernel<{vt}>",
             "fft_fwd": f"fft_fwd_kernel<{vt}>", "reduce": f"reduce_kernel<{vt}>"}.get(dom, dom)
    if dom in ("gemm_translation", "gemm_solve"):
        peak = PEAK["fp32"] if (mixed and dom == "gemm_translation") else PEAK["fp64"]
        roof = {"bound": "mfma", "achieved": per[dom]["work"] / t_dom / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s"}
        work_desc = "useful complex-GEMM flops 8*M*K*N over the class's groups"
    elif dom == "kapply":
        work = per[dom]["work"] / nrhs * (20 + 8 * nrhs)
        peak = PEAK["fp32"] if mixed else PEAK["fp64"]
        roof = {"bound": "mfma", "achieved": work / t_dom / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s"}
        work_desc = "target-source interactions x (20 + 8*nrhs) flops (VALU)"
    elif dom == "reduce":
        work = per[dom]["work"] * esz
        roof = {"bound": "hbm", "achieved": work / t_dom / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s"}
        work_desc = "contribution slots read (targets written not counted)"
    elif dom == "vlist":
        work = ((st["v_sources"] + st["v_targets"]) * 1728 * esz * nrhs + st["nlevels_lf"] * 316 * 1728 * esz)
        roof = {"bound": "hbm", "achieved": work / t_dom / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s"}
        work_desc = "compulsory bytes: (V sources + V targets) x 1728 spectra entries + translation tables"
    else:
        roof = {"bound": "hbm", "achieved": 0.0, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s"}
        work_desc = "n/a"
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    tfile = os.path.join(ROOT, "profiles", f"{args.workload}_{args.precision}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roof["traffic"] = json.load(f).get(kname)
    roof["kernel"] = kname
    roof["class"] = dom
    roof["launches_per_eval"] = launches
    roof["avg_launch_us"] = t_dom / max(launches, 1) * 1e6
    roof["work"] = work_desc
    hip.free(dd)
    hip.free(vd)
    w.close()

    serve = None
    if args.serve:
        log("serving leg ...")
        try:
            serve = serve_bench(args.serve.split(","), args.serve_qps, args.serve_requests, 16, args.precision)
            log(f"serving: p50 {serve['p50_ms']:.2f} ms p99 {serve['p99_ms']:.2f} ms, "
                f"{serve['throughput_M_point_evals_per_s']:.1f} M point-evals/s, {serve['batches']} batches")
        except Exception as e:  # the serving leg must not hide the main metric
            serve = {"error": repr(e)}
            log(f"serving leg failed: {e!r}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(mlib_gen.library_paths(os.path.join(ROOT, "ddfmm_amd", "_ops"), args.kernel, 2, 4),
                           args.kernel)
        log(f"cpu baseline {cpu['value']:.3f} M/s on {cpu['cores']} cores")

    value = N * nrhs / dt / 1e6
    out = {
        "metric": "M point-evals/s·RHS at rel.err<=1e-5 vs FP64 direct",
        "value": value,
        "unit": "M point-evals/s·RHS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "fp32+fp64 (FP32: translations with the HF solves folded in, M2L, near field, V list; FP64: LF check-to-equivalent solves)", "fp64": "fp64",
                  "fp64_fused": "fp64 (fused solves)"}[args.precision],
        "data": "synthetic",
        "config": {"workload": args.workload, "n_points": N, "nrhs": nrhs, "kernel": "exp(ikr)/r" if args.kernel == 0
                   else "exp(ikr)", "K": wl["K"], "ACCU": 2, "NPQ": 4, "ptsmax": wl["ptsmax"],
                   "precision": args.precision, "rel_err_vs_direct": err, "n_check": int(len(chk)),
                   "meets_tolerance": bool(err <= 1e-5)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "serve": serve,
        "per_kernel_ms": {k: round(v["ms"], 4) for k, v in per.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
