"""Writes the golden fixtures in this directory (inputs and expected outputs).

Expected outputs come from the CPU oracle and are cross-checked here against
independent numpy implementations before anything is written; the script
refuses to write a fixture on any disagreement.

  sort_<kt>.npz     edge-count inputs (0, 1, 2, 7, 63, 64, 65, 257, 1000, 3841)
                    with many ties; stable-sorted keys and original
                    indices for less, greater and radix-bit order
  merge_u32.npz     sorted pairs (A, B) with ties and empty sides; merged keys
                    and source indices (ties take A)
  scan.npz          int32 inputs incl. wrap-around, exclusive/inclusive scans
                    (uint32 modular) and float32 inputs with exact float64
                    prefix sums
  kat.json          sha256 digests for the reference examples' inputs
                    (glibc rand(), default seed): vk_radix2, vk_sort, vk_merge,
                    gl_sort

Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

SIZES = [0, 1, 2, 7, 63, 64, 65, 257, 1000, 3841]
NP = oracle.NP


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def keys_for(kt, n, seed):
    kind = {"u32": "lowentropy", "i32": "u32", "u64": "u64", "f32": "f32_pm1"}[kt]
    x = oracle.generate(n, kind, seed)
    if kt == "i32":
        x = x.view(np.int32)
    if n > 10:
        x = x[np.arange(n) % (n // 3 + 1)]  # ties
    return x.copy()


def radix_bits(x, kt):
    if kt in ("u32", "u64"):
        return x
    if kt == "i32":
        return x.view(np.uint32) ^ np.uint32(0x80000000)
    b = x.view(np.uint32)
    return b ^ np.where(b >> np.uint32(31), np.uint32(0xFFFFFFFF), np.uint32(0x80000000))


def desc_order(x, kt):
    if kt in ("u32", "u64"):
        return np.argsort(~x, kind="stable")
    return np.argsort(-x.astype(np.float64), kind="stable")


def make_sort(kt):
    out = {}
    for i, n in enumerate(SIZES):
        x = keys_for(kt, n, 100 + i)
        v = np.arange(n, dtype=np.uint32)
        for tag, (k, vv), order in [
            ("less", oracle.stable_sort(x, v, kt), np.argsort(x, kind="stable")),
            ("greater", oracle.stable_sort(x, v, kt, greater=True), desc_order(x, kt)),
            ("radix", oracle.stable_sort_radix(x, v, kt), np.argsort(radix_bits(x, kt), kind="stable")),
        ]:
            assert np.array_equal(vv, order.astype(np.uint32)), (kt, n, tag)
            assert np.array_equal(k.view(np.uint8), x[order].view(np.uint8)), (kt, n, tag)
            out["%d_%s" % (n, tag)] = vv
        out["%d_in" % n] = x
    np.savez_compressed(os.path.join(HERE, "sort_%s.npz" % kt), **out)


def make_merge():
    out = {}
    cases = [(0, 0), (0, 5), (5, 0), (1, 1), (100, 37), (3840, 3841), (5000, 9000)]
    for i, (na, nb) in enumerate(cases):
        a = np.sort(oracle.generate(na, "u32", 200 + i) % 1000)
        b = np.sort(oracle.generate(nb, "u32", 300 + i) % 1000)
        av = np.arange(na, dtype=np.uint32)
        bv = np.arange(na, na + nb, dtype=np.uint32)
        c, cv = oracle.merge(a, b, av, bv)
        order = np.argsort(np.concatenate([a, b]), kind="stable")
        assert np.array_equal(cv, order.astype(np.uint32))
        out["%d_%d_a" % (na, nb)] = a
        out["%d_%d_b" % (na, nb)] = b
        out["%d_%d_src" % (na, nb)] = cv
    np.savez_compressed(os.path.join(HERE, "merge_u32.npz"), **out)


def make_scan():
    out = {}
    for i, n in enumerate([1, 5, 4095, 4097]):
        x = oracle.generate(n, "i32_0_16", 400 + i)
        with np.errstate(over="ignore"):
            inc = (np.cumsum(x.astype(np.uint64)) & 0xFFFFFFFF).astype(np.uint32)
        exc = np.concatenate([np.zeros(1, np.uint32), inc[:-1]])
        assert np.array_equal(oracle.exclusive_scan_u32(x), exc)
        assert np.array_equal(oracle.inclusive_scan_u32(x), inc)
        out["i32_%d_in" % n] = x
        out["i32_%d_exc" % n] = exc
        out["i32_%d_inc" % n] = inc
        f = oracle.generate(n, "f32_01", 500 + i)
        exact = np.cumsum((f.astype(np.float64) * 2 ** 24).astype(np.int64)).astype(np.float64) * 2.0 ** -24
        assert np.array_equal(oracle.inclusive_scan_f64(f), exact)
        out["f32_%d_in" % n] = f
        out["f32_%d_inc" % n] = exact
    wrap = np.full(3000, 2 ** 30 + 3, dtype=np.int32)
    out["wrap_in"] = wrap
    out["wrap_exc"] = oracle.exclusive_scan_u32(wrap)
    np.savez_compressed(os.path.join(HERE, "scan.npz"), **out)


def make_kat():
    n = 256 * 32 * 32 * 32
    r = oracle.glibc_rand(n)
    s = np.sort(r)
    assert np.array_equal(s, oracle.std_sort_u32(r))
    r1 = oracle.glibc_rand(20000)
    vk_sort_keys = (r1[:10000] % 100000).astype(np.float32)
    a = np.sort((r1[:10000] % 100000).astype(np.float32))
    b = np.sort((r1[10000:] % 100000).astype(np.float32))
    gl = (r1[:10000] % 10000).astype(np.float32)
    kat = {
        "vk_radix2": {"source": "examples/vk_radix2/vk_radix2.cxx:14-75", "count": n,
                      "input_sha256": sha(r), "sorted_sha256": sha(s)},
        "vk_sort": {"source": "examples/vk_sort/vk_sort.cxx:13-16", "count": 10000,
                    "sorted_sha256": sha(np.sort(vk_sort_keys))},
        "vk_merge": {"source": "examples/vk_merge/vk_merge.cxx:13-27", "count": 20000,
                     "merged_sha256": sha(np.sort(np.concatenate([a, b])))},
        "gl_sort": {"source": "examples/gl_sort/gl_sort.cxx:30-42", "count": 10000,
                    "indices_sha256": sha(np.argsort(gl, kind="stable").astype(np.int32)),
                    "sorted_sha256": sha(np.sort(gl))},
        "vk_scan": {"source": "examples/vk_scan/vk_scan.cxx:17-23", "count": 10000,
                    "closed_form": "out[i] = i*(i-1)/2"},
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    for kt in ("u32", "i32", "u64", "f32"):
        make_sort(kt)
    make_merge()
    make_scan()
    make_kat()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
