"""Cross-check against an independent implementation: AMD's rocPRIM
(vendor_rocprim.so, compiled from ROCm's source headers by the build) run on
the same device inputs.  Stable sorts, the stable merge and integer scans have
a unique result, so mgpu-hip's output must equal rocPRIM's bit for bit, at
sizes well beyond what the CPU oracle checks element by element.  Skipped
when the comparator library was not built."""
import ctypes
import os

import numpy as np
import pytest
import torch

import mgpu_hip
from gpu_util import banner, gen_dev
from mgpu_hip import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "vendor_rocprim.so")

vp, i64 = ctypes.c_void_p, ctypes.c_int64
sz = ctypes.POINTER(ctypes.c_size_t)
pint = ctypes.POINTER(ctypes.c_int)
SIGS = {"vendor_radix_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, vp, pint],
        "vendor_radix_sort_pairs_u64": [vp, sz, vp, vp, vp, vp, i64, vp, pint],
        "vendor_merge_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, vp],
        "vendor_merge_u32": [vp, sz, vp, i64, vp, i64, vp, vp],
        "vendor_exclusive_scan_i32": [vp, sz, vp, vp, i64, ctypes.c_int32, vp],
        "vendor_exclusive_scan_u64": [vp, sz, vp, vp, i64, ctypes.c_uint64, vp],
        "vendor_segmented_radix_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, i64, vp, vp, vp]}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        pytest.skip("vendor_rocprim.so not built")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGS.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int
    return L


def p(t):
    return vp(t.data_ptr())


def call(fn, *args, tail=()):
    """rocPRIM's two-phase call on the current stream, then synchronize."""
    stream = vp(torch.cuda.current_stream().cuda_stream)
    need = ctypes.c_size_t(0)
    assert fn(None, ctypes.byref(need), *args, stream, *tail) == 0
    temp = torch.empty(max(need.value, 1), dtype=torch.uint8, device="cuda")
    cap = ctypes.c_size_t(temp.numel())
    assert fn(p(temp), ctypes.byref(cap), *args, stream, *tail) == 0
    torch.cuda.synchronize()


def same(a, b):
    return bool(torch.equal(a.view(torch.int32), b.view(torch.int32)))


@pytest.mark.parametrize("n", [(1 << 24) + 77, 1 << 26])
@pytest.mark.parametrize("kind", ["u32", "lowentropy"])
def test_radix_and_mergesort_pairs_u32(lib, n, kind):
    banner("rocPRIM cross-check: u32 pairs", n, kind)
    keys = gen_dev(n, kind, 11)
    vals = gen_dev(n, "iota", 0)
    k, v = keys.clone(), vals.clone()
    mgpu_hip.radix_sort(k, v)
    k2, v2 = keys.clone(), vals.clone()
    kb, vb = torch.empty_like(keys), torch.empty_like(vals)
    alt = ctypes.c_int(0)
    call(lib.vendor_radix_sort_pairs_u32, p(k2), p(kb), p(v2), p(vb), n, tail=(ctypes.byref(alt),))
    rk, rv = (kb, vb) if alt.value else (k2, v2)
    assert same(k, rk) and same(v, rv), "radix sort differs from rocPRIM"
    k3, v3 = keys.clone(), vals.clone()
    mgpu_hip.mergesort_kv(k3, v3)
    assert same(k3, k) and same(v3, v), "mergesort differs from radix sort"
    call(lib.vendor_merge_sort_pairs_u32, p(keys), p(kb), p(vals), p(vb), n)
    assert same(k3, kb) and same(v3, vb), "mergesort differs from rocPRIM"


def test_radix_pairs_u64(lib):
    n = (1 << 24) + 5
    banner("rocPRIM cross-check: u64 pairs", n)
    keys = gen_dev(n, "u64", 12)
    vals = gen_dev(n, "iota", 0)
    k, v = keys.clone(), vals.clone()
    mgpu_hip.radix_sort(k, v)
    kb, vb = torch.empty_like(keys), torch.empty_like(vals)
    alt = ctypes.c_int(0)
    call(lib.vendor_radix_sort_pairs_u64, p(keys), p(kb), p(vals), p(vb), n,
         tail=(ctypes.byref(alt),))
    rk, rv = (kb, vb) if alt.value else (keys, vals)
    assert same(k, rk) and same(v, rv), "u64 radix sort differs from rocPRIM"


def test_merge_and_scans(lib):
    m = (1 << 25) + 3
    banner("rocPRIM cross-check: merge, scans", m)
    a, b = gen_dev(m, "lowentropy", 13), gen_dev(m - 7, "lowentropy", 14)
    mgpu_hip.radix_sort(a)
    mgpu_hip.radix_sort(b)
    c, c2 = torch.empty(2 * m - 7, dtype=torch.uint32, device="cuda"), \
        torch.empty(2 * m - 7, dtype=torch.uint32, device="cuda")
    mgpu_hip.merge_keys(a, b, c, "less")
    call(lib.vendor_merge_u32, p(a), m, p(b), m - 7, p(c2))
    assert same(c, c2), "merge differs from rocPRIM"
    s = (1 << 27) + 9
    x = gen_dev(s, "i32_0_16", 15)
    y, y2 = torch.empty_like(x), torch.empty_like(x)
    mgpu_hip.scan_into(x, y, init=7)
    call(lib.vendor_exclusive_scan_i32, p(x), p(y2), s, 7)
    assert same(y, y2), "int32 scan differs from rocPRIM"
    x64 = x[: 2 * ((s - 1) // 2)].view(torch.uint64)
    y64, y64b = torch.empty_like(x64), torch.empty_like(x64)
    mgpu_hip.scan_into(x64, y64, init=3)
    call(lib.vendor_exclusive_scan_u64, p(x64), p(y64b), x64.numel(), 3)
    assert same(y64, y64b), "uint64 scan differs from rocPRIM"


def test_segmented_sort_pairs(lib):
    lens = workloads.lognormal_lengths(20000, 512.0, 1.5, 1, 1 << 18, seed=3)
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    banner("rocPRIM cross-check: segmented sort", n, len(lens))
    segs = torch.from_numpy(starts).to("cuda")
    ends = torch.from_numpy(np.append(starts[1:], np.int32(n)).astype(np.int32)).to("cuda")
    keys = gen_dev(n, "lowentropy", 16)
    vals = gen_dev(n, "iota", 0)
    k, v = keys.clone(), vals.clone()
    mgpu_hip.segmented_sort(k, segs, v)
    kb, vb = torch.empty_like(keys), torch.empty_like(vals)
    call(lib.vendor_segmented_radix_sort_pairs_u32, p(keys), p(kb), p(vals), p(vb), n,
         len(starts), p(segs), p(ends))
    assert same(k, kb) and same(v, vb), "segmented sort differs from rocPRIM"
