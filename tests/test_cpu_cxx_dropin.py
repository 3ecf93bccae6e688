"""The header-only C++ drop-in (include/ddfmm_wave3d.hpp) compiles with g++ against the
C ABI library and drives a host-only setup with the reference's wiring."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cxx_dropin_builds_and_plans(oplib, tmp_path):
    exe = tmp_path / "dropin"
    lib_dir = os.path.join(ROOT, "ddfmm_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "dropin_main.cpp"), "-L", lib_dir, "-lddfmm",
                    f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), oplib[0], oplib[1]], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "npts 3000" in out.stdout
    assert "not set up on a device" in out.stdout
