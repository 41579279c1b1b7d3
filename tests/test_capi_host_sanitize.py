"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only, no GPU): every csrc/*.hip translation unit is compiled with the
sanitizers on the host side only (device code is built as usual and never
runs), linked with tests/cpp/capi_host_sanitize.cpp, and run.  The driver
sizes every entry point at edge counts and checks argument validation,
limits, settings and status strings; a sanitizer report fails the test."""
import concurrent.futures
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mgpu-shaders_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Sanitizers for the host side of each translation unit only.
HOST_SANITIZERS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
FLAGS = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-fno-omit-frame-pointer",
         "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "mgpu-shaders_amd", "include"),
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + HOST_SANITIZERS


def _compile(src, obj, extra=()):
    r = subprocess.run([HIPCC] + FLAGS + list(extra) + ["-c", src, "-o", obj],
                       capture_output=True, text=True, timeout=1200)
    return src, r.returncode, r.stderr[-3000:]


def test_capi_host_code_under_sanitizers(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is needed for the sanitized build")
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = [(os.path.join(CSRC, f), str(tmp_path / (f[:-4] + ".o")), ()) for f in srcs]
    jobs.append((os.path.join(ROOT, "tests", "cpp", "capi_host_sanitize.cpp"),
                 str(tmp_path / "driver.o"), ("-x", "hip")))
    workers = max(1, min(len(jobs), os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for src, rc, err in pool.map(lambda j: _compile(*j), jobs):
            assert rc == 0, "%s:\n%s" % (src, err)
    exe = str(tmp_path / "capi_host")
    r = subprocess.run([HIPCC] + FLAGS + [o for _, o, _ in jobs] + ["-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ALL PASS" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
