"""The oracle reproduces every committed golden fixture (CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = [0, 1, 2, 7, 63, 64, 65, 257, 1000, 3841]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("kt", ["u32", "i32", "u64", "f32"])
def test_sort_fixtures(kt):
    z = np.load(os.path.join(G, "sort_%s.npz" % kt))
    for n in SIZES:
        x = z["%d_in" % n]
        v = np.arange(n, dtype=np.uint32)
        assert np.array_equal(oracle.stable_sort(x, v, kt)[1], z["%d_less" % n])
        assert np.array_equal(oracle.stable_sort(x, v, kt, greater=True)[1], z["%d_greater" % n])
        assert np.array_equal(oracle.stable_sort_radix(x, v, kt)[1], z["%d_radix" % n])
        assert np.array_equal(oracle.mergesort_schedule(x, v, 4, 3, kt)[1], z["%d_less" % n])
        assert np.array_equal(oracle.radix_schedule(x, v, 64, kt)[1], z["%d_radix" % n])


def test_merge_fixtures():
    z = np.load(os.path.join(G, "merge_u32.npz"))
    for key in [k for k in z.files if k.endswith("_src")]:
        stem = key[:-4]
        a, b = z[stem + "_a"], z[stem + "_b"]
        av = np.arange(len(a), dtype=np.uint32)
        bv = np.arange(len(a), len(a) + len(b), dtype=np.uint32)
        assert np.array_equal(oracle.merge(a, b, av, bv)[1], z[key])


def test_scan_fixtures():
    z = np.load(os.path.join(G, "scan.npz"))
    for n in [1, 5, 4095, 4097]:
        x = z["i32_%d_in" % n]
        assert np.array_equal(oracle.exclusive_scan_u32(x), z["i32_%d_exc" % n])
        assert np.array_equal(oracle.inclusive_scan_u32(x), z["i32_%d_inc" % n])
        assert np.array_equal(oracle.scan_schedule_u32(x, 896), z["i32_%d_exc" % n])
        assert np.array_equal(oracle.inclusive_scan_f64(z["f32_%d_in" % n]), z["f32_%d_inc" % n])
    assert np.array_equal(oracle.exclusive_scan_u32(z["wrap_in"]), z["wrap_exc"])


def test_kat_digests():
    kat = json.load(open(os.path.join(G, "kat.json")))
    n = kat["vk_radix2"]["count"]
    r = oracle.glibc_rand(n)
    assert sha(r) == kat["vk_radix2"]["input_sha256"]
    assert sha(oracle.std_sort_u32(r)) == kat["vk_radix2"]["sorted_sha256"]
    r1 = oracle.glibc_rand(20000)
    gl = (r1[:10000] % 10000).astype(np.float32)
    k, idx = oracle.mergesort_schedule(gl, None, 128, 7, "f32", sort_indices=True)
    assert sha(idx.astype(np.int32)) == kat["gl_sort"]["indices_sha256"]
    assert sha(k) == kat["gl_sort"]["sorted_sha256"]
