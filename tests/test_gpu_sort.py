"""GPU parity: mergesort, merge and radix sort against the oracle, bit-exact."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, dev_equal, gen_dev, to_dev, to_host

pytestmark = pytest.mark.gpu

KTS = ("u32", "i32", "u64", "f32", "i64", "f64")
MS_NV = 512 * 19      # mergesort tile of 32-bit keys in the C ABI instantiation
MS_NV64 = 512 * 13    # mergesort tile of 64-bit keys
MERGE_NV64 = 512 * 15  # merge tile of 64-bit keys
MERGE_NV = 512 * 23   # merge tile of 32-bit keys
RADIX_NV = 512 * 8    # radix tile of every size below 2^22 keys
SIZES = [1, 2, 7, 64, 100, MS_NV64 - 1, MS_NV64, MS_NV64 + 1, MS_NV - 1, MS_NV, MS_NV + 1,
         2 * MS_NV + 3, 5 * MS_NV - 11, RADIX_NV + 1, 100003, (1 << 20) + 17]
# Tile counts whose merge levels end in partial groups of four runs (a last
# group holding one, two or three runs, some of them short).
GROUP_SIZES = [9 * MS_NV - 5, 17 * MS_NV + 2, 33 * MS_NV, 50 * MS_NV + 77, 130 * MS_NV - 1]


def make_keys(kt, n, seed, pattern="random"):
    kind = {"u32": "u32", "i32": "u32", "u64": "u64", "f32": "f32_pm1", "i64": "u64",
            "f64": "u64"}[kt]
    if pattern == "lowentropy" and kt in ("u32", "i32"):
        kind = "lowentropy"
    x = oracle.wide_keys(kt, oracle.generate(n, kind, seed))
    if pattern == "ties":
        x = x[np.arange(n) % 97] if n else x
    elif pattern == "equal":
        x = np.full(n, x[0] if n else 0, dtype=x.dtype)
    elif pattern == "sorted":
        x = np.sort(x)
    elif pattern == "reversed":
        x = np.sort(x)[::-1].copy()
    return x


def dev_keys(x):
    return to_dev(x)


@pytest.fixture(params=[mgpu_hip.RADIX_COUNT_SCAN_SCATTER, mgpu_hip.RADIX_SINGLE_PASS],
                ids=["count_scan_scatter", "single_pass"])
def radix_algo(request):
    """Runs the test under each digit-pass structure of the radix sort."""
    old = mgpu_hip.radix_algorithm()
    mgpu_hip.set_radix_algorithm(request.param)
    yield request.param
    mgpu_hip.set_radix_algorithm(old)


@pytest.fixture(params=[mgpu_hip.MERGESORT_TWO_WAY, mgpu_hip.MERGESORT_FOUR_WAY],
                ids=["two_way", "four_way"])
def merge_passes(request):
    """Runs the test under each merge pass structure (at these sizes the
    automatic choice would always be two-way)."""
    old = mgpu_hip.mergesort_passes()
    mgpu_hip.set_mergesort_passes(request.param)
    yield request.param
    mgpu_hip.set_mergesort_passes(old)


# ------------------------------------------------------------------ mergesort

@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kt", KTS)
def test_mergesort_kv(n, kt, merge_passes):
    banner("mergesort_kv", kt, n, merge_passes)
    keys = make_keys(kt, n, 1, "ties")
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.mergesort_kv(dk, dv)
    ek, ev = oracle.stable_sort(keys, vals, kt)
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


@pytest.mark.parametrize("kt,n", [("u64", (1 << 24) + 5), ("u32", (24 << 20) + 3),
                                  ("u32", (20 << 20) - 7)])
def test_mergesort_kv_automatic_structure(kt, n):
    """The automatic pass structure around its key-value threshold (168 MiB
    per pass: four-way for the first two sizes, two-way for the last), against
    the oracle's stable sort."""
    banner("mergesort_kv automatic", kt, n)
    old = mgpu_hip.mergesort_passes()
    try:
        mgpu_hip.set_mergesort_passes(0)
        keys = make_keys(kt, n, 7, "ties")
        vals = np.arange(n, dtype=np.uint32)
        dk, dv = dev_keys(keys), to_dev(vals)
        mgpu_hip.mergesort_kv(dk, dv)
    finally:
        mgpu_hip.set_mergesort_passes(old)
    ek, ev = oracle.stable_sort(keys, vals, kt)
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


@pytest.mark.parametrize("n", GROUP_SIZES)
@pytest.mark.parametrize("kt", ["u32", "f64"])
@pytest.mark.parametrize("comp", ["less", "greater"])
def test_mergesort_kv_partial_groups(n, kt, comp, merge_passes):
    banner("mergesort_kv groups", kt, n, comp, merge_passes)
    keys = make_keys(kt, n, 5, "ties" if comp == "less" else "random")
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.mergesort_kv(dk, dv, comp=comp)
    ek, ev = oracle.stable_sort(keys, vals, kt, greater=(comp == "greater"))
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


@pytest.mark.parametrize("pattern", ["random", "equal", "sorted", "reversed", "lowentropy"])
@pytest.mark.parametrize("comp", ["less", "greater"])
def test_mergesort_patterns(pattern, comp, merge_passes):
    n = 3 * MS_NV + 1001
    banner("mergesort pattern", pattern, comp, merge_passes)
    keys = make_keys("u32", n, 2, pattern)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.mergesort_kv(dk, dv, comp=comp)
    ek, ev = oracle.stable_sort(keys, vals, "u32", greater=(comp == "greater"))
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("comp", ["less", "greater"])
def test_mergesort_keys(kt, comp, merge_passes):
    n = 200001
    banner("mergesort_keys", kt, comp, merge_passes)
    keys = make_keys(kt, n, 3)
    dk = dev_keys(keys)
    mgpu_hip.mergesort_keys(dk, comp=comp)
    np.testing.assert_array_equal(to_host(dk), oracle.stable_sort(keys, kt=kt, greater=(comp == "greater")))


@pytest.mark.parametrize("kt", ["f32", "u64"])
@pytest.mark.parametrize("comp", ["less", "greater"])
def test_mergesort_indices(kt, comp, merge_passes):
    """indices[i] = i before the sort, then sorted with the keys: the stable
    argsort, over enough tiles for two four-way passes."""
    n = 100003
    banner("mergesort_indices", kt, comp, merge_passes)
    keys = make_keys(kt, n, 8, "ties")
    dk = dev_keys(keys)
    di = torch.empty(n, dtype=torch.int32, device="cuda")
    mgpu_hip.mergesort_indices(dk, di, comp=comp)
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.stable_sort(keys, vals, kt, greater=(comp == "greater"))
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(di).view(np.uint32), ev)


def test_mergesort_indices_gl_sort_known_answer():
    # examples/gl_sort/gl_sort.cxx:33-42: 10000 floats rand() % 10000 sorted
    # with indices; gathered == sorted and the output is sorted.
    banner("gl_sort kat")
    keys = (oracle.glibc_rand(10000) % 10000).astype(np.float32)
    dk = dev_keys(keys)
    di = torch.empty(10000, dtype=torch.int32, device="cuda")
    mgpu_hip.mergesort_indices(dk, di)
    k, idx = to_host(dk), to_host(di)
    np.testing.assert_array_equal(keys[idx], k)
    assert np.all(k[:-1] <= k[1:])
    np.testing.assert_array_equal(idx, np.argsort(keys, kind="stable"))


def test_mergesort_vk_sort_known_answer():
    # examples/vk_sort/vk_sort.cxx:15-16: 10000 floats rand() % 100000.
    banner("vk_sort kat")
    keys = (oracle.glibc_rand(10000) % 100000).astype(np.float32)
    dk = dev_keys(keys)
    mgpu_hip.mergesort_keys(dk)
    np.testing.assert_array_equal(to_host(dk), np.sort(keys))


# ---------------------------------------------------------------------- merge

MERGE_SIZES = [(0, 1), (1, 0), (1, 1), (5, 3000), (3000, 5), (MS_NV64, MS_NV64),
               (MERGE_NV64 // 2, MERGE_NV64 // 2 + 1), (MERGE_NV64, 1),
               (MERGE_NV // 2, MERGE_NV // 2), (MERGE_NV, 1), (10007, 20011),
               (1 << 19, (1 << 19) + 3)]


@pytest.mark.parametrize("na,nb", MERGE_SIZES)
@pytest.mark.parametrize("kt", KTS)
def test_merge_kv(na, nb, kt):
    banner("merge_kv", kt, na, nb)
    a = oracle.stable_sort(make_keys(kt, na, 4, "ties"), kt=kt)
    b = oracle.stable_sort(make_keys(kt, nb, 5, "ties"), kt=kt)
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    da, db, dav, dbv = dev_keys(a), dev_keys(b), to_dev(av), to_dev(bv)
    dc = torch.empty(na + nb, dtype=da.dtype, device="cuda")
    dcv = torch.empty(na + nb, dtype=torch.uint32, device="cuda")
    mgpu_hip.merge(da, dav, db, dbv, dc, dcv, "less")
    ec, ecv = oracle.merge(a, b, av, bv, kt)
    np.testing.assert_array_equal(to_host(dc), ec)
    np.testing.assert_array_equal(to_host(dcv), ecv)


@pytest.mark.parametrize("comp", ["less", "greater"])
def test_merge_keys_all_equal_and_greater(comp):
    banner("merge_keys", comp)
    na, nb = 50000, 70001
    greater = comp == "greater"
    a = oracle.stable_sort(make_keys("u32", na, 6, "ties"), greater=greater)
    b = oracle.stable_sort(make_keys("u32", nb, 7, "ties"), greater=greater)
    da, db = dev_keys(a), dev_keys(b)
    dc = torch.empty(na + nb, dtype=da.dtype, device="cuda")
    mgpu_hip.merge_keys(da, db, dc, comp)
    np.testing.assert_array_equal(to_host(dc), oracle.merge(a, b, greater=greater))


@pytest.mark.parametrize("order", ["a_below_b", "b_below_a", "b_inside_a"])
@pytest.mark.parametrize("comp", ["less", "greater"])
@pytest.mark.parametrize("kt", ["u32", "u64"])
def test_merge_kv_disjoint_ranges(order, comp, kt):
    """Merge-path co-ranks on the edges of the grid: every A key below every B
    key (and the reverse), and B falling inside one run of equal A keys; the
    arrays span many tiles, so every partition lies on an edge."""
    banner("merge disjoint", order, comp, kt)
    na, nb = 300007, 200003
    dt = np.uint64 if kt == "u64" else np.uint32

    def draw(n, seed, base):
        return oracle.generate(n, "u32", seed).astype(dt) % 1000 + dt(base)
    if order == "a_below_b":
        a, b = draw(na, 40, 0), draw(nb, 41, 1000)
    elif order == "b_below_a":
        a, b = draw(na, 40, 1000), draw(nb, 41, 0)
    else:  # A: runs of 500, 700 and 900; B: all 700, inside A's run of 700s
        a = np.concatenate([np.full(na // 3, 500), np.full(na // 3, 700),
                            np.full(na - 2 * (na // 3), 900)]).astype(dt)
        b = np.full(nb, 700, dtype=dt)
    greater = comp == "greater"
    a = oracle.stable_sort(a, kt=kt, greater=greater)
    b = oracle.stable_sort(b, kt=kt, greater=greater)
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    dc = torch.empty(na + nb, dtype=dev_keys(a).dtype, device="cuda")
    dcv = torch.empty(na + nb, dtype=torch.uint32, device="cuda")
    mgpu_hip.merge(dev_keys(a), to_dev(av), dev_keys(b), to_dev(bv), dc, dcv, comp)
    ec, ecv = oracle.merge(a, b, av, bv, kt, greater=greater)
    np.testing.assert_array_equal(to_host(dc), ec)
    np.testing.assert_array_equal(to_host(dcv), ecv)


def test_merge_vk_merge_known_answer():
    # examples/vk_merge/vk_merge.cxx:18-27: two sorted arrays of 10000 floats
    # rand() % 100000 (a first, then b from the same rand() stream).
    banner("vk_merge kat")
    r = oracle.glibc_rand(20000) % 100000
    a = np.sort(r[:10000].astype(np.float32))
    b = np.sort(r[10000:].astype(np.float32))
    dc = torch.empty(20000, dtype=torch.float32, device="cuda")
    mgpu_hip.merge_keys(dev_keys(a), dev_keys(b), dc, "less")
    np.testing.assert_array_equal(to_host(dc), np.sort(np.concatenate([a, b])))


# ---------------------------------------------------------------------- radix

@pytest.mark.parametrize("n", SIZES + [3 * RADIX_NV, 3 * RADIX_NV + 5])
@pytest.mark.parametrize("kt", KTS)
def test_radix_kv(n, kt, radix_algo):
    banner("radix_kv", kt, n, radix_algo)
    keys = make_keys(kt, n, 8, "ties")
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.radix_sort(dk, dv)
    ek, ev = oracle.stable_sort_radix(keys, vals, kt)
    np.testing.assert_array_equal(to_host(dk).view(np.uint8), ek.view(np.uint8))
    np.testing.assert_array_equal(to_host(dv), ev)


# Sorts of 4097 to 4178 tiles of the large-sort shapes (6144 keys per tile for
# 64-bit keys, 10240 / 12288 for 32-bit key-value / keys), each ending in a
# partial tile: the 64-bit passes 1-7 count the digit bytes the previous
# scatter wrote, the 32-bit passes the keys.
LARGE_TILES = [("u64", 4178 * 256 * 24 - 5196, True), ("f64", 4097 * 256 * 24 - 1, True),
               ("u32", 4101 * 512 * 20 - 77, True), ("u32", 4097 * 512 * 24 - 1, False),
               ("u64", 4099 * 512 * 12 - 3, False)]


@pytest.mark.parametrize("kt,n,kv", LARGE_TILES)
def test_radix_large_tile_counts(kt, n, kv):
    banner("radix large tile counts", kt, n, "kv" if kv else "keys")
    keys = make_keys(kt, n, 13)
    dk = dev_keys(keys)
    if kv:
        vals = np.arange(n, dtype=np.uint32)
        dv = to_dev(vals)
        mgpu_hip.radix_sort(dk, dv)
        ek, ev = oracle.stable_sort_radix(keys, vals, kt)
        np.testing.assert_array_equal(to_host(dv), ev)
    else:
        mgpu_hip.radix_sort(dk)
        ek = oracle.stable_sort_radix(keys, None, kt)
    np.testing.assert_array_equal(to_host(dk).view(np.uint8), ek.view(np.uint8))


@pytest.mark.parametrize("pattern", ["random", "equal", "sorted", "reversed", "lowentropy"])
@pytest.mark.parametrize("kt", ["u32", "u64"])
def test_radix_patterns(pattern, kt, radix_algo):
    n = 5 * RADIX_NV + 77
    banner("radix pattern", pattern, kt, radix_algo)
    keys = make_keys(kt, n, 9, pattern)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.radix_sort(dk, dv)
    ek, ev = oracle.stable_sort_radix(keys, vals, kt)
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


def test_radix_special_floats(radix_algo):
    banner("radix special floats", radix_algo)
    base = np.array([np.nan, np.inf, 1.0, 0.0, -0.0, -1.0, -np.inf, 1e-45, -1e-45, 3.4e38,
                     -3.4e38], dtype=np.float32)
    neg_nan = np.array([0xFFC00000], dtype=np.uint32).view(np.float32)
    keys = np.tile(np.concatenate([base, neg_nan]), 3000)
    vals = np.arange(len(keys), dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.radix_sort(dk, dv)
    ek, ev = oracle.stable_sort_radix(keys, vals, "f32")
    np.testing.assert_array_equal(to_host(dk).view(np.uint32), ek.view(np.uint32))
    np.testing.assert_array_equal(to_host(dv), ev)


def test_radix_special_doubles_and_int64_extremes(radix_algo):
    banner("radix special doubles, int64 extremes", radix_algo)
    base = np.array([np.nan, np.inf, 1.0, 0.0, -0.0, -1.0, -np.inf, 5e-324, -5e-324, 1.7e308,
                     -1.7e308], dtype=np.float64)
    neg_nan = np.array([0xFFF8000000000000], dtype=np.uint64).view(np.float64)
    keys = np.tile(np.concatenate([base, neg_nan]), 2000)
    vals = np.arange(len(keys), dtype=np.uint32)
    dk, dv = dev_keys(keys), to_dev(vals)
    mgpu_hip.radix_sort(dk, dv)
    ek, ev = oracle.stable_sort_radix(keys, vals, "f64")
    np.testing.assert_array_equal(to_host(dk).view(np.uint64), ek.view(np.uint64))
    np.testing.assert_array_equal(to_host(dv), ev)
    ints = np.tile(np.array([-2**63, -2**32, -1, 0, 1, 2**32, 2**63 - 1], dtype=np.int64), 5000)
    ints = ints[oracle.generate(len(ints), "u32", 3).argsort(kind="stable")]
    vals = np.arange(len(ints), dtype=np.uint32)
    dk, dv = dev_keys(ints), to_dev(vals)
    mgpu_hip.radix_sort(dk, dv)
    ek, ev = oracle.stable_sort_radix(ints, vals, "i64")
    np.testing.assert_array_equal(to_host(dk), ek)
    np.testing.assert_array_equal(to_host(dv), ev)


def test_radix_keys_vk_radix2_known_answer(radix_algo):
    # examples/vk_radix2/vk_radix2.cxx:14-75: 8,388,608 glibc rand() keys,
    # sorted six times in place, compared with std::sort.
    banner("vk_radix2 kat")
    n = 256 * 32 * 32 * 32
    keys = oracle.glibc_rand(n)
    dk = dev_keys(keys)
    aux = None
    for _ in range(6):
        aux = mgpu_hip.radix_sort(dk, aux=aux)
    np.testing.assert_array_equal(to_host(dk), oracle.std_sort_u32(keys))


def test_radix_repeat_is_deterministic_and_idempotent(radix_algo):
    banner("radix repeat", radix_algo)
    n = (1 << 21) + 5
    keys = gen_dev(n, "lowentropy", 10)
    vals = gen_dev(n, "iota")
    k1, v1 = keys.clone(), vals.clone()
    mgpu_hip.radix_sort(k1, v1)
    k2, v2 = keys.clone(), vals.clone()
    mgpu_hip.radix_sort(k2, v2)
    assert dev_equal(k1, k2) and dev_equal(v1, v2)
    k3, v3 = k1.clone(), v1.clone()
    mgpu_hip.radix_sort(k3, v3)
    assert dev_equal(k3, k1) and dev_equal(v3, v1)


@pytest.mark.parametrize("kt,kv,n", [
    ("u32", True, 6 * (1 << 20)),             # last count on 512 x 8 tiles (kv)
    ("u32", True, 615 * 10240 + 1),           # 512 x 20 tiles, one key past a tile edge
    ("f32", False, 1 << 22),                  # last count on 512 x 8 tiles (keys)
    ("i32", False, 683 * 6144),               # 512 x 12 tiles, exactly whole tiles
    ("u32", False, 2731 * 12288 + 5),         # 512 x 24 tiles
    ("u64", True, 683 * 6144 + 3),            # 64-bit keys on 512 x 12 tiles
])
def test_radix_tile_shape_boundaries(kt, kv, n, radix_algo):
    """Sizes on either side of the counts where the library changes tile shape."""
    banner("radix shape boundary", kt, kv, n, radix_algo)
    keys = make_keys(kt, n, 21, "ties")
    dk = dev_keys(keys)
    if kv:
        vals = np.arange(n, dtype=np.uint32)
        dv = to_dev(vals)
        mgpu_hip.radix_sort(dk, dv)
        ek, ev = oracle.stable_sort_radix(keys, vals, kt)
        np.testing.assert_array_equal(to_host(dv), ev)
    else:
        mgpu_hip.radix_sort(dk)
        ek, _ = oracle.stable_sort_radix(keys, np.zeros(n, np.uint32), kt)
    np.testing.assert_array_equal(to_host(dk).view(np.uint8), ek.view(np.uint8))


# ------------------------------------------------------ look-back fallback path

@pytest.mark.parametrize("limit", [0, 1])
@pytest.mark.parametrize("kt", ["u32", "u64", "f32"])
def test_radix_forced_lookback_fallback(limit, kt, radix_algo):
    """A tiny spin limit makes tiles count late predecessors from their input
    (the scan's tile sums, or the single pass's per-digit counts); the sort
    must be unchanged."""
    banner("radix fallback", kt, limit, radix_algo)
    n = 40 * RADIX_NV + 123
    keys = make_keys(kt, n, 12, "ties")
    vals = np.arange(n, dtype=np.uint32)
    old = mgpu_hip.lookback_spin_limit()
    mgpu_hip.lookback_fallbacks(reset=True)
    try:
        mgpu_hip.set_lookback_spin_limit(limit)
        dk, dv = dev_keys(keys), to_dev(vals)
        mgpu_hip.radix_sort(dk, dv)
        torch.cuda.synchronize()
    finally:
        mgpu_hip.set_lookback_spin_limit(old)
    print("fallbacks", mgpu_hip.lookback_fallbacks(), flush=True)
    ek, ev = oracle.stable_sort_radix(keys, vals, kt)
    np.testing.assert_array_equal(to_host(dk).view(np.uint8), ek.view(np.uint8))
    np.testing.assert_array_equal(to_host(dv), ev)
