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
