"""Build libmgpu_hip.so (gfx950) from csrc/*.hip with hipcc, in-tree.

Usage: python mgpu-shaders_amd/build.py [-v]

Each translation unit is compiled in parallel into _build/, then linked into
mgpu_hip/libmgpu_hip.so next to the Python bindings.  Packed-FP32 instruction
selection is disabled for the device code (-packed-fp32-ops).
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
OUT = os.path.join(PKG, "mgpu_hip", "libmgpu_hip.so")
# C++ consumers of the header templates, built as test executables.
CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
CPP_BIN = os.path.join(CPP_TESTS, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CXXFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "include"),
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CXXFLAGS + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    lines = [l for l in (r.stdout + r.stderr).splitlines()
             if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, "\n".join(lines)))
    if verbose and lines:
        print("\n".join(lines))
    return obj


def _cpp_test(src, verbose):
    exe = os.path.join(CPP_BIN, os.path.basename(src)[:-4])
    cmd = [HIPCC] + CXXFLAGS + [src, "-o", exe]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        lines = [l for l in (r.stdout + r.stderr).splitlines()
                 if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, "\n".join(lines)))
    return exe


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(CPP_BIN, exist_ok=True)
    srcs = sources()
    tests = sorted(os.path.join(CPP_TESTS, f) for f in os.listdir(CPP_TESTS) if f.endswith(".hip"))
    workers = min(len(srcs) + len(tests), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        test_jobs = [ex.submit(_cpp_test, t, verbose) for t in tests]
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
        for j in test_jobs:
            j.result()
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
