"""GPU outputs equal the committed golden fixtures and known answers."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = [0, 1, 2, 7, 63, 64, 65, 257, 1000, 3841]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("kt", ["u32", "i32", "u64", "f32"])
def test_sort_fixtures_on_gpu(kt):
    z = np.load(os.path.join(G, "sort_%s.npz" % kt))
    for n in SIZES:
        banner("golden sort", kt, n)
        x = z["%d_in" % n]
        for algo, comp, tag in [("merge", "less", "less"), ("merge", "greater", "greater"),
                                ("radix", None, "radix")]:
            dk = to_dev(x)
            dv = to_dev(np.arange(n, dtype=np.uint32))
            if n:
                if algo == "merge":
                    mgpu_hip.mergesort_kv(dk, dv, comp=comp)
                else:
                    mgpu_hip.radix_sort(dk, dv)
            src = z["%d_%s" % (n, tag)]
            np.testing.assert_array_equal(to_host(dv), src)
            np.testing.assert_array_equal(to_host(dk).view(np.uint8), x[src].view(np.uint8))


def test_merge_fixtures_on_gpu():
    z = np.load(os.path.join(G, "merge_u32.npz"))
    for key in [k for k in z.files if k.endswith("_src")]:
        stem = key[:-4]
        banner("golden merge", stem)
        a, b = z[stem + "_a"], z[stem + "_b"]
        na, nb = len(a), len(b)
        if na + nb == 0:
            continue
        dc = torch.empty(na + nb, dtype=torch.uint32, device="cuda")
        dcv = torch.empty(na + nb, dtype=torch.uint32, device="cuda")
        mgpu_hip.merge(to_dev(a), to_dev(np.arange(na, dtype=np.uint32)), to_dev(b),
                       to_dev(np.arange(na, na + nb, dtype=np.uint32)), dc, dcv, "less")
        np.testing.assert_array_equal(to_host(dcv), z[key])
        np.testing.assert_array_equal(to_host(dc), np.concatenate([a, b])[z[key]])


def test_scan_fixtures_on_gpu():
    z = np.load(os.path.join(G, "scan.npz"))
    for n in [1, 5, 4095, 4097]:
        banner("golden scan", n)
        x = to_dev(z["i32_%d_in" % n])
        out = torch.empty_like(x)
        mgpu_hip.scan_into(x, out)
        np.testing.assert_array_equal(to_host(out).view(np.uint32), z["i32_%d_exc" % n])
        mgpu_hip.scan_into(x, out, inclusive=True)
        np.testing.assert_array_equal(to_host(out).view(np.uint32), z["i32_%d_inc" % n])
        f = to_dev(z["f32_%d_in" % n])
        fo = torch.empty_like(f)
        mgpu_hip.scan_into(f, fo, inclusive=True)
        exact = z["f32_%d_inc" % n]
        got = to_host(fo).astype(np.float64)
        assert np.max(np.abs(got - exact) / exact) <= 2e-6
    w = to_dev(z["wrap_in"])
    mgpu_hip.scan(w)
    np.testing.assert_array_equal(to_host(w).view(np.uint32), z["wrap_exc"])


def test_kat_digests_on_gpu():
    kat = json.load(open(os.path.join(G, "kat.json")))
    banner("kat gl_sort / vk_sort / vk_merge")
    r1 = oracle.glibc_rand(20000)
    gl = to_dev((r1[:10000] % 10000).astype(np.float32))
    idx = torch.empty(10000, dtype=torch.int32, device="cuda")
    mgpu_hip.mergesort_indices(gl, idx)
    assert sha(to_host(idx)) == kat["gl_sort"]["indices_sha256"]
    assert sha(to_host(gl)) == kat["gl_sort"]["sorted_sha256"]
    vs = to_dev((r1[:10000] % 100000).astype(np.float32))
    mgpu_hip.mergesort_keys(vs)
    assert sha(to_host(vs)) == kat["vk_sort"]["sorted_sha256"]
    a = to_dev(np.sort((r1[:10000] % 100000).astype(np.float32)))
    b = to_dev(np.sort((r1[10000:] % 100000).astype(np.float32)))
    c = torch.empty(20000, dtype=torch.float32, device="cuda")
    mgpu_hip.merge_keys(a, b, c, "less")
    assert sha(to_host(c)) == kat["vk_merge"]["merged_sha256"]
    banner("kat vk_radix2")
    k = to_dev(oracle.glibc_rand(kat["vk_radix2"]["count"]))
    aux = None
    for _ in range(6):
        aux = mgpu_hip.radix_sort(k, aux=aux)
    assert sha(to_host(k)) == kat["vk_radix2"]["sorted_sha256"]
