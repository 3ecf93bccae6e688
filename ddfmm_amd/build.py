"""Build the in-tree HIP library ``ddfmm_amd/libddfmm.so`` for gfx950 with hipcc.

One shared object holds the host planner (C++), the HIP kernels and the C ABI of
include/ddfmm.h.  Packed-FP32 instruction selection is disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libddfmm.so")
SOURCES = ["tree.cpp", "mlib.cpp", "planner.cpp", "tiles.cpp", "capi.cpp", "kernels.hip"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 5) -> str:
    """Compile every source (in parallel) and link libddfmm.so; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              # MFMA accumulators in ordinary VGPRs: with the AGPR form the GEMM's
              # loop-carried accumulators are copied between register files every slab
              "-mllvm", "-amdgpu-mfma-vgpr-form=1",
              # only the C ABI of include/ddfmm.h is exported: callers that include the
              # C++ drop-in headers see none of the library's internal C++ symbols
              "-fvisibility=hidden", "-fvisibility-inlines-hidden",
              "-Wno-unused-result", "-I", CSRC]
    procs, objs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmd = [cc] + common + lang + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len([p for p in procs if p[1].poll() is None]) >= jobs:
            procs[0][1].wait()
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {s}")
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "--offload-arch=gfx950", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
