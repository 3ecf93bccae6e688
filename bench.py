#!/usr/bin/env python3
"""Throughput of the directional-FMM evaluation pass on one MI355X.

One step = one Wave3d::eval over the workload (densities and potentials resident in
HBM, eval_device).  Default workload: the largest single-GPU configuration, the
perturbed ellipsoid with semi-axes 32 x 8 x 8 wavelengths (radial factor
1 + 0.1 sin(3 theta) cos(5 phi)), N = 4,000,000 area-uniform points (seed 3), 8
complex-normal density vectors, kernel exp(ikr)/r, K = 128, 8^3 cells, ptsmax = 400,
ACCU=2 (P=6), NPQ=4, mixed precision.  Metric: M point-evaluations/s per RHS, reported
with the relative L2 error at 1000 random targets against an FP64 direct sum.

The same JSON line carries the plain-FP64 time of the same workload (the mixed/FP64
ratio), the 1M-point sphere (1 RHS) and the 250k-point sphere (64 RHS) in both
precisions, the roofline of the dominant kernel (HIP events on the stages' streams), a
serving leg (p50/p99 latency) and a CPU baseline on a bounded sample.
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

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp64_fused": 2, "mixed_split": 3}   # ddfmm.h precision modes
ST_ZERO, ST_GATHER, ST_SCATTER, ST_KAPPLY, ST_GEMM, ST_FFTFWD, ST_VLIST, ST_REDUCE, ST_FORK, ST_JOIN = range(10)


def kernel_class(kind: int, prec: int) -> str:
    if kind == ST_GEMM:
        return "gemm_solve" if prec == 1 else "gemm_translation"
    return {ST_KAPPLY: "kapply", ST_VLIST: "vlist", ST_FFTFWD: "fft_fwd", ST_ZERO: "zero",
            ST_GATHER: "gather", ST_SCATTER: "scatter", ST_REDUCE: "reduce", ST_FORK: "sync", ST_JOIN: "sync"}[kind]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ops_paths, kernel, workload="ellipsoid_4m_x8"):
    """The compiled FP64 restatement of the FMM (oracle/wave3d_ref.cpp, OpenMP) on the host
    cores: one right-hand side of `workload` (the reference's eval is single-RHS), setup
    excluded, eval timed; threads = the cores this process may use, at most 16."""
    from ddfmm_amd import geometry, serialize
    from oracle import wave3d_cref, wave3d_ref

    cores = min(16, len(os.sched_getaffinity(0)))
    wl = geometry.WORKLOADS[workload]
    pos = geometry.workload_points(workload)
    den = geometry.densities(len(pos), wl["seed"], nrhs=1)
    ld, hd = serialize.read_ld(ops_paths[0]), serialize.read_hd(ops_paths[1])
    t0 = time.perf_counter()
    c = wave3d_cref.Wave3dCOracle(pos, wl["K"], ld, hd, ktype=kernel, ptsmax=wl["ptsmax"], maxlevel=14,
                                  numC=wl.get("numC", 1)).setup()
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    val = c.eval(den, threads=cores)
    dt = time.perf_counter() - t0
    c.close()
    chk = geometry.check_targets(len(pos), 200, wl["seed"])
    err = wave3d_ref.relerr(val[chk], wave3d_ref.direct_sum(pos[chk], pos, den, kernel, threads=cores))
    log(f"cpu baseline: {workload} 1 RHS on {cores} threads: eval {dt:.1f}s (setup {setup_s:.1f}s), "
        f"rel.err vs direct {err:.2e}")
    return {"value": len(pos) / dt / 1e6, "unit": "M point-evals/s·RHS", "cores": cores, "kind": "port",
            "sample": f"compiled FP64 OpenMP restatement of Wave3d::setup+eval (oracle/wave3d_ref.cpp) on "
                      f"{workload} ({len(pos)} points, K={wl['K']:g}), one RHS, eval only: {dt:.1f} s "
                      f"(setup {setup_s:.1f} s), rel.err {err:.1e} vs FP64 direct at 200 targets",
            "workload": workload, "eval_s": dt}


def serve_bench(names, qps, count, slot, precision, seed=4):
    """Poisson request stream over resident geometries (continuous batching)."""
    from ddfmm_amd import geometry, mlib_gen
    from ddfmm_amd.serve import GpuEngine, Server, latency_stats, poisson_requests

    ops = mlib_gen.ensure_library(os.path.join(ROOT, "ddfmm_amd", "_ops"), ktype=0, accu=2, NPQ=4,
                                  hf_max_W=4, lf_min_W=2.0 ** -3)
    pts = {n: geometry.workload_points(n) for n in names}

    def make(key):
        wl = geometry.WORKLOADS[key]
        return GpuEngine(key, pts[key], {"-wave3d_ACCU": 2, "-wave3d_K": wl["K"], "-wave3d_ptsmax": wl["ptsmax"],
                                         "-wave3d_maxlevel": 14}, ops, slot,
                         PRECISIONS[precision],
                         numC=wl.get("numC", 1))

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


def _vlist_name(vt, nrhs):
    """The V-list accumulation instance the launcher picks for nrhs (kernels.hip)."""
    if nrhs == 1:
        return f"vlist_acc_kernel<{vt}>"
    r2 = 2 if vt == "float2" else 1
    f, r1 = (1, 8) if nrhs >= 8 * r2 else (2, 4) if nrhs >= 4 * r2 else (4, 2)
    return f"vlist_acc_multi_kernel<{vt},{f},{r1}>"


# kernel family names as the kernel trace prints them (dominant class -> instance)
KNAME = {
    "gemm_translation": lambda vt, n: "gemm_kernel<float2,float2,float>" if vt == "float2"
    else "gemm_kernel<double2,double2,double>",
    "gemm_solve": lambda vt, n: "gemm_kernel<double2,float2,double>" if vt == "float2"
    else "gemm_kernel<double2,double2,double>",
    "kapply": lambda vt, n: f"kapply_kernel<{vt},{rhs_block(n)}>",
    "kapply_mfma": lambda nb, kt: f"kapply_mfma_kernel<{nb},{kt},false>",
    "vlist": _vlist_name,
    "fft_fwd": lambda vt, n: f"fft_fwd_kernel<{vt}>" if n == 1 else
    f"fft_fwd_multi_kernel<{vt},{8 if vt == 'float2' else 4},2>",
    "reduce": lambda vt, n: f"reduce_kernel<{vt}>",
}


def rhs_block(nrhs: int) -> int:
    """Right-hand sides one kernel-apply / V-list block carries (as the launcher picks it)."""
    return 1 if nrhs == 1 else 4 if nrhs <= 4 else 8


def measure(workload, precision, kernel, nrhs, steps, warmup, nchk, detail=False, calib=False, tune=True,
            graph=True):
    """Set up one Wave3d on the workload, time `steps` evaluations (after `warmup`) with
    densities and potentials resident in HBM, and check 1000 targets against the FP64
    direct sum.  With detail, also per-stage HIP-event times and the dominant kernel's
    roofline."""
    from ddfmm_amd import geometry, mlib_gen
    from ddfmm_amd._lib import Hip
    from ddfmm_amd.wave3d import Kernel3d, Mlib3d, Wave3d, direct_sum

    wl = dict(geometry.WORKLOADS[workload])
    nrhs = nrhs or wl["nrhs"]
    ops = mlib_gen.ensure_library(os.path.join(ROOT, "ddfmm_amd", "_ops"), ktype=kernel, accu=2, NPQ=4,
                                  hf_max_W=4, lf_min_W=2.0 ** -3)
    t0 = time.perf_counter()
    pos = geometry.workload_points(workload)
    den = geometry.densities(len(pos), wl["seed"], nrhs=nrhs)
    N = len(pos)
    log(f"{workload} {precision}: N={N} nrhs={nrhs} kernel={kernel} ({time.perf_counter() - t0:.1f}s)")
    ml = Mlib3d()
    ml.setup({"-mlib3d_ldname": ops[0], "-mlib3d_hdname": ops[1]})
    w = Wave3d()
    w.posptr, w.mlibptr, w.kernel = pos, ml, Kernel3d(kernel)
    w.geomprtn = np.zeros((wl.get("numC", 1),) * 3, dtype=np.int32)
    t0 = time.perf_counter()
    w.setup({"-wave3d_ACCU": 2, "-wave3d_NPQ": 4, "-wave3d_K": wl["K"], "-wave3d_ptsmax": wl["ptsmax"],
             "-wave3d_maxlevel": 14}, nrhs_max=nrhs,
            precision=PRECISIONS[precision])
    setup_s = time.perf_counter() - t0
    st = w.stats()
    log(f"setup {setup_s:.1f}s: leaves={st['nleaf']} m2l_pairs={st['m2l_pairs']} "
        f"v_pairs={st['v_pairs']} device={st['device_bytes'] / 2**30:.2f} GiB")
    tuned = None
    if tune:
        t0 = time.perf_counter()
        tuned = w.autotune(nrhs, reps=1)
        tuned["seconds"] = time.perf_counter() - t0
        log(f"autotune {tuned['seconds']:.1f}s: GEMM stages {tuned['gemm_ms_planner']:.3f} ms (planner tiles) -> "
            f"{tuned['gemm_ms_tuned']:.3f} ms (tuned), heights {tuned['tiles']}")
    if graph:
        w.set_graph(True)
    hip = Hip()
    if calib:
        nb = 256 << 20
        a, b = hip.malloc(nb), hip.malloc(nb)
        hip.copy_d2d(b, a, nb)
        hip.sync()
        hip.free(a)
        hip.free(b)
    den2 = np.ascontiguousarray(np.atleast_2d(den))
    dd = hip.to_device(den2)
    vd = hip.malloc(den2.nbytes)
    for _ in range(warmup):
        w.eval_device(dd, vd, nrhs)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        w.eval_device(dd, vd, nrhs)
    hip.sync()
    dt = (time.perf_counter() - t0) / steps
    log(f"{steps} steps: {dt * 1e3:.3f} ms/step")
    val = hip.to_host(vd, np.empty_like(den2))
    chk = geometry.check_targets(N, nchk, wl["seed"])
    exact = direct_sum(pos[chk], pos, den2, kernel)
    err = float(np.linalg.norm(val[:, chk] - exact) / np.linalg.norm(exact))
    log(f"rel.err vs FP64 direct at {len(chk)} targets: {err:.3e}")
    out = {"workload": workload, "precision": precision, "n_points": N, "nrhs": nrhs, "K": wl["K"],
           "ptsmax": wl["ptsmax"], "kernel": "exp(ikr)/r" if kernel == 0 else "exp(ikr)",
           "ms_per_step": dt * 1e3, "value": N * nrhs / dt / 1e6, "rel_err_vs_direct": err,
           "n_check": int(len(chk)), "setup_s": setup_s, "device_GiB": st["device_bytes"] / 2**30,
           "autotune": tuned, "graph": bool(graph)}
    if detail:
        out.update(stage_detail(w, st, dd, vd, nrhs, precision, workload, kernel))
    hip.free(dd)
    hip.free(vd)
    w.close()
    return out


def stage_detail(w, st, dd, vd, nrhs, precision, workload, kernel=0):
    """Per-stage HIP-event times on the stages' own streams, the kernel classes and the
    roofline of the dominant class (work per launch / average launch duration)."""
    w.set_timing(True)
    per, per_stage = {}, {}
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
        sorted(per_stage.items(), key=lambda kv: -kv[1][0])[:14]))
    log("per kernel class (ms/eval, launches): " + ", ".join(
        f"{k}={v['ms']:.3f}/{v['launches']}" for k, v in sorted(per.items(), key=lambda kv: -kv[1]["ms"])))
    dom = max(per, key=lambda k: per[k]["ms"])
    t_dom = per[dom]["ms"] * 1e-3
    launches = per[dom]["launches"]
    mixed = precision in ("mixed", "mixed_split")
    split = precision == "mixed_split"
    esz = 8 if mixed else 16
    vt = "float2" if mixed else "double2"
    kname = KNAME.get(dom, lambda vt, n: dom)(vt, nrhs)
    kmfma = mixed and nrhs >= 4         # contexts of >= 4 RHS apply kernels on the matrix cores
    if dom == "kapply" and kmfma:
        kname = KNAME["kapply_mfma"](2 if nrhs > 8 else 1, kernel)
    if dom in ("gemm_translation", "gemm_solve"):
        peak = PEAK["fp32"] if (mixed and dom == "gemm_translation") else PEAK["fp64"]
        work_desc = "useful complex-GEMM flops 8*M*K*N over the class's groups"
        if split and dom == "gemm_translation":
            # six bf16 products per FP32 product: the useful-flop ceiling is the bf16 peak / 6
            peak = PEAK["bf16"] / 6
            kname = "gemm_bf16s_kernel"
            work_desc += " (split-BF16: 6 bf16 MFMA products each; peak = dense bf16 peak / 6)"
        roof = {"bound": "mfma", "achieved": per[dom]["work"] / t_dom / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s"}
    elif dom == "kapply":
        work = per[dom]["work"] / nrhs * (20 + 8 * nrhs)
        peak = PEAK["fp32"] if mixed else PEAK["fp64"]
        roof = {"bound": "mfma" if kmfma else "valu", "achieved": work / t_dom / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s"}
        work_desc = ("target-source interactions x (20 + 8*nrhs) flops: G on the vector ALU, the "
                     "8*nrhs products on the matrix cores (FP32)" if kmfma else
                     "target-source interactions x (20 + 8*nrhs) flops on the vector ALU")
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
    tfile = os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roof["traffic"] = json.load(f).get(kname)
    roof["kernel"] = kname
    roof["class"] = dom
    roof["launches_per_eval"] = launches
    roof["avg_launch_us"] = t_dom / max(launches, 1) * 1e6
    roof["work"] = work_desc
    # the same stages one at a time on one stream: each time is the stage's own, and the
    # dominant class's launches run with nothing beside them (frac_alone)
    w.set_timing(2)
    serial = {}
    t_alone = 0.0
    for _ in range(reps):
        w.eval_device(dd, vd, nrhs)
        for name, ms, work, kind, prec in w.stage_times():
            if kind in (ST_FORK, ST_JOIN):
                continue
            serial[name] = serial.get(name, 0.0) + ms / reps
            if kernel_class(kind, prec) == dom:
                t_alone += ms * 1e-3 / reps
    w.set_timing(False)
    if t_alone > 0:
        roof["achieved_alone"] = roof["achieved"] * t_dom / t_alone     # same work, its own time
        roof["frac_alone"] = roof["achieved_alone"] / roof["peak"]
        roof["avg_launch_us_alone"] = t_alone / max(launches, 1) * 1e6
    log("stages alone (ms): " + ", ".join(f"{n}={v:.3f}" for n, v in sorted(serial.items(), key=lambda kv: -kv[1])[:14])
        + f"; sum {sum(serial.values()):.2f}")
    return {"roofline": roof, "per_kernel_ms": {k: round(v["ms"], 4) for k, v in per.items()},
            "per_stage_ms": {k: round(v[0], 4) for k, v in per_stage.items()},
            "per_stage_alone_ms": {k: round(v, 4) for k, v in serial.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ellipsoid_4m_x8",
                    help="headline workload (default: the largest single-GPU configuration)")
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--nrhs", type=int, default=None)
    ap.add_argument("--precision", default="mixed", choices=sorted(PRECISIONS))
    ap.add_argument("--no-fp64-leg", action="store_true",
                    help="skip the plain-FP64 run of the headline workload (the mixed/FP64 ratio)")
    ap.add_argument("--extra", default="sphere32_1m,sphere16_250k_x64,sphere16_250k_x64@1",
                    help="comma-separated further workloads timed in mixed and plain FP64, '<name>@1' for the "
                         "kernel exp(ikr) ('' to skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="one 256 MiB device-to-device copy (blit kernel) to calibrate PMC byte counters")
    ap.add_argument("--chk", type=int, default=1000)
    ap.add_argument("--no-autotune", action="store_true", help="keep the planner's GEMM tiling")
    ap.add_argument("--no-graph", action="store_true", help="issue the stages one by one (no graph replay)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-workload", default="ellipsoid_4m_x8",
                    help="workload of the CPU baseline (one RHS, eval timed)")
    ap.add_argument("--serve", default="sphere8_60k,sphere16_250k_x64,sphere32_1m,ellipsoid_4m_x8",
                    help="comma-separated geometries for the serving leg ('' to skip)")
    ap.add_argument("--serve-qps", type=float, default=10.0)
    ap.add_argument("--serve-requests", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(one device share per process, no collectives).", file=sys.stderr)
        sys.exit(2)

    from ddfmm_amd import mlib_gen

    tg = dict(tune=not args.no_autotune, graph=not args.no_graph)
    head = measure(args.workload, args.precision, args.kernel, args.nrhs, args.steps, args.warmup, args.chk,
                   detail=True, calib=args.calib_copy, **tg)
    fp64 = None
    if not args.no_fp64_leg and args.precision != "fp64":
        fp64 = measure(args.workload, "fp64", args.kernel, args.nrhs, max(2, args.steps // 2), 1, args.chk, **tg)
    extra = {}
    for item in [x for x in args.extra.split(",") if x and x != args.workload]:
        wname, _, kt = item.partition("@")      # "<workload>@1": kernel exp(ikr)
        kt = int(kt) if kt else args.kernel
        m = measure(wname, "mixed", kt, None, max(2, args.steps // 2), 1, args.chk, **tg)
        f = measure(wname, "fp64", kt, None, max(2, args.steps // 2), 1, args.chk, **tg)
        extra[item] = {"kernel": m["kernel"], "nrhs": m["nrhs"], "n_points": m["n_points"], "mixed_ms": m["ms_per_step"],
                        "mixed_value": m["value"], "mixed_rel_err": m["rel_err_vs_direct"],
                        "fp64_ms": f["ms_per_step"], "fp64_value": f["value"], "fp64_rel_err": f["rel_err_vs_direct"],
                        "mixed_vs_fp64": f["ms_per_step"] / m["ms_per_step"]}

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
                           args.kernel, args.cpu_workload)
        same = head if args.cpu_workload == args.workload else \
            {"value": extra.get(args.cpu_workload, {}).get("mixed_value")}
        if same.get("value"):
            cpu["gpu_mixed_value_same_workload"] = same["value"]
        log(f"cpu baseline {cpu['value']:.3f} M/s on {cpu['cores']} cores")

    err = head["rel_err_vs_direct"]
    cfg = {"workload": args.workload, "n_points": head["n_points"], "nrhs": head["nrhs"],
           "kernel": head["kernel"], "K": head["K"], "ACCU": 2, "NPQ": 4, "ptsmax": head["ptsmax"],
           "precision": args.precision, "rel_err_vs_direct": err, "n_check": head["n_check"],
           "meets_tolerance": bool(err <= 1e-5), "device_GiB": round(head["device_GiB"], 2),
           "graph_replay": head["graph"], "autotune": head["autotune"]}
    if fp64 is not None:
        cfg.update({"fp64_ms_per_step": fp64["ms_per_step"], "fp64_value": fp64["value"],
                    "fp64_rel_err_vs_direct": fp64["rel_err_vs_direct"],
                    "mixed_vs_fp64": fp64["ms_per_step"] / head["ms_per_step"]})
    out = {
        "metric": "M point-evals/s·RHS at rel.err<=1e-5 vs FP64 direct",
        "value": head["value"],
        "unit": "M point-evals/s·RHS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "fp32+fp64 (FP32: translations with the HF solves folded in, M2L, near field, V list; "
                           "FP64: LF check-to-equivalent solves, coordinate differences)", "fp64": "fp64",
                  "fp64_fused": "fp64 (fused solves)",
                  "mixed_split": "bf16x3+fp32+fp64 (translation GEMMs as three-term split-BF16 products "
                                 "accumulated in FP32; otherwise as mixed)"}[args.precision],
        "data": "synthetic",
        "config": cfg,
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "serve": serve,
        "configs": extra,
        "per_kernel_ms": head["per_kernel_ms"],
        "per_stage_ms": head["per_stage_ms"],
        "per_stage_alone_ms": head["per_stage_alone_ms"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
