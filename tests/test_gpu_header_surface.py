"""GPU: the C++ header-template surface as a consumer uses it (user
comparators through mergesort and merge, launch / transform lambdas,
transform_scan with a functor), run as the executable built from
tests/cpp/header_surface.hip."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "header_surface")


def test_header_surface_program():
    if not os.path.exists(EXE):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % EXE)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout


def test_capture_before_first_call_program():
    """A template caller that begins a global-mode graph capture before any
    library call (tests/cpp/capture_first.hip, a fresh process): nothing the
    library does on first use may allocate or synchronise inside the
    capture, and three replays match the std:: algorithms."""
    exe = os.path.join(os.path.dirname(EXE), "capture_first")
    if not os.path.exists(exe):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout


def test_context_port_program():
    """A call site in the reference's host idiom (context_t, cmd_buffer_t,
    memcache_t, int merge-path partitions) ported to mgpu/hip
    (tests/cpp/context_port.hip)."""
    exe = os.path.join(os.path.dirname(EXE), "context_port")
    if not os.path.exists(exe):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout
