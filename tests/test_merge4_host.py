"""Four-way merge partitions checked on the host (no GPU): the device search
functions of mgpu/hip/merge.hxx and mgpu/hip/mergesort.hxx are host-callable,
so tests/cpp/merge4_host.cpp compiles them with hipcc as host code and checks
them against a brute-force stable four-way merge at every diagonal."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_merge4_partitions_against_brute_force(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc is needed to compile the host check")
    exe = str(tmp_path / "merge4_host")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "mgpu-shaders_amd", "include"),
           os.path.join(ROOT, "tests", "cpp", "merge4_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ALL PASS" in r.stdout
