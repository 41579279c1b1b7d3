"""Pins the CPU oracle before it is trusted as the checker (CPU only).

Every oracle function is checked against something independent of it: numpy
implementations, closed forms, brute force on tiny inputs, and the known
answers of SplitMix64 and glibc rand().  Planted mistakes (oracle.planted)
show that each check rejects a plausible error.
"""
import numpy as np
import pytest

import oracle
from mgpu_hip import workloads

KTS = ("u32", "i32", "u64", "f32", "i64", "f64")


def np_splitmix_word(seed, stream, i):
    """Independent numpy restatement of the counter-based SplitMix64 word."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(40)) + (np.uint64(stream) << np.uint64(36)) \
            + i.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(32)).astype(np.uint32)


def keys_for(kt, n, seed, low_entropy=False):
    kind = {"u32": "u32", "i32": "u32", "u64": "u64", "f32": "f32_pm1", "i64": "u64",
            "f64": "u64"}[kt]
    if low_entropy and kt in ("u32", "i32"):
        kind = "lowentropy"
    return oracle.wide_keys(kt, oracle.generate(n, kind, seed))


def np_stable_order(keys, kt, greater=False):
    if not greater:
        return np.argsort(keys, kind="stable")
    if kt in ("u32", "u64", "i32", "i64"):
        return np.argsort(~np_radix_bits(keys, kt), kind="stable")
    return np.argsort(-(keys.astype(np.float64)), kind="stable")


def np_radix_bits(keys, kt):
    if kt in ("u32", "u64"):
        return keys
    if kt == "i32":
        return keys.view(np.uint32) ^ np.uint32(0x80000000)
    if kt == "i64":
        return keys.view(np.uint64) ^ np.uint64(1 << 63)
    if kt == "f64":
        b = keys.view(np.uint64)
        mask = np.where(b >> np.uint64(63), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(1 << 63))
        return b ^ mask
    b = keys.view(np.uint32)
    mask = np.where(b >> np.uint32(31), np.uint32(0xFFFFFFFF), np.uint32(0x80000000))
    return b ^ mask


# ------------------------------------------------------------------ inputs

def test_splitmix_known_answer():
    # First output of SplitMix64 seeded with 0 (Steele, Lea & Flood 2014;
    # java.util.SplittableRandom): 0xE220A8397B1DCDAF.
    assert oracle.generate(1, "u32", 0)[0] == 0xE220A839


@pytest.mark.parametrize("kind", ["u32", "lowentropy", "i32_0_16", "f32_01", "u64", "f32_pm1"])
def test_generators_match_numpy(kind):
    n, seed, off = 5000, 3, 12345
    got = oracle.generate(n, kind, seed, off)
    i = np.arange(off, off + n, dtype=np.uint64)
    w = [np_splitmix_word(seed, s, i) for s in range(4)]
    if kind == "u32":
        exp = w[0]
    elif kind == "lowentropy":
        exp = w[0] & w[1] & w[2] & w[3]
    elif kind == "i32_0_16":
        exp = (w[0] >> np.uint32(28)).astype(np.int32)
    elif kind == "f32_01":
        exp = ((w[0] >> np.uint32(8)).astype(np.float64) * 2.0 ** -24).astype(np.float32)
    elif kind == "u64":
        exp = (w[0].astype(np.uint64) << np.uint64(32)) | w[1].astype(np.uint64)
    else:
        exp = ((w[0] >> np.uint32(8)).astype(np.float64) * 2.0 ** -23 - 1.0).astype(np.float32)
    np.testing.assert_array_equal(got, exp)


def test_generator_ranges():
    f = oracle.generate(100000, "f32_01")
    assert f.min() >= 0 and f.max() < 1
    g = oracle.generate(100000, "f32_pm1")
    assert g.min() >= -1 and g.max() < 1
    h = oracle.generate(100000, "i32_0_16")
    assert h.min() == 0 and h.max() == 15
    le = oracle.generate(1 << 16, "lowentropy")
    # P(low byte == 0) for an AND of four uniform bytes = (15/16)^8 ~ 0.597.
    p0 = np.mean((le & 0xFF) == 0)
    assert 0.57 < p0 < 0.62


def test_glibc_rand_known_answer():
    # glibc rand() with the default seed 1.
    assert list(oracle.glibc_rand(5)) == [1804289383, 846930886, 1681692777, 1714636915,
                                          1957747793]


# ----------------------------------------------------------- std oracles

@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("greater", [False, True])
@pytest.mark.parametrize("low", [False, True])
def test_stable_sort_vs_numpy(kt, greater, low):
    n = 20000
    keys = keys_for(kt, n, 1, low)
    if not low:
        keys = keys[np.arange(n) % 997]  # many ties
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.stable_sort(keys, vals, kt, greater)
    order = np_stable_order(keys, kt, greater)
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("kt", KTS)
def test_stable_sort_radix_vs_numpy(kt):
    n = 20000
    keys = keys_for(kt, n, 2)[np.arange(n) % 1999]
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.stable_sort_radix(keys, vals, kt)
    order = np.argsort(np_radix_bits(keys, kt), kind="stable")
    np.testing.assert_array_equal(k.view(np.uint8), keys[order].view(np.uint8))
    np.testing.assert_array_equal(v, vals[order])


def test_radix_bits_special_floats():
    specials = np.array([np.nan, np.inf, 1.0, 0.0, -0.0, -1.0, -np.inf,
                         np.float32(1e-45), -np.float32(1e-45), 3.4e38, -3.4e38],
                        dtype=np.float32)
    neg_nan = np.array([0xFFC00000], dtype=np.uint32).view(np.float32)
    x = np.concatenate([specials, neg_nan])
    bits = oracle.radix_bits(x, "f32")
    np.testing.assert_array_equal(bits, np_radix_bits(x, "f32"))
    back = oracle.radix_bits(bits, "f32", inverse=True)
    np.testing.assert_array_equal(back.view(np.uint32), x.view(np.uint32))
    order = x[np.argsort(bits, kind="stable")]
    names = [repr(float(v)) if not np.isnan(v) else ("-nan" if np.signbit(v) else "nan")
             for v in order]
    assert names[0] == "-nan" and names[-1] == "nan"
    finite = order[1:-1]
    assert np.all(finite[:-1] <= finite[1:])
    # -0.0 sorts strictly before +0.0 in radix order.
    iz = [i for i, v in enumerate(order) if v == 0]
    assert np.signbit(order[iz[0]]) and not np.signbit(order[iz[1]])


def test_radix_bits_signed_and_u64():
    i = np.array([-2**31, -1, 0, 1, 2**31 - 1], dtype=np.int32)
    b = oracle.radix_bits(i, "i32")
    assert np.all(b[:-1] < b[1:])
    u = np.array([0, 1, 2**63, 2**64 - 1], dtype=np.uint64)
    np.testing.assert_array_equal(oracle.radix_bits(u, "u64"), u)
    # 64-bit signed keys: the sign flip is a 64-bit shift (the reference's
    # int-literal shift, cta_radix.hxx:18,44, would be undefined here).
    j = np.array([-2**63, -2**32, -1, 0, 1, 2**32, 2**63 - 1], dtype=np.int64)
    bj = oracle.radix_bits(j, "i64")
    assert np.all(bj[:-1] < bj[1:])
    np.testing.assert_array_equal(bj, np_radix_bits(j, "i64"))
    np.testing.assert_array_equal(oracle.radix_bits(bj, "i64", inverse=True), j)


def test_radix_bits_special_doubles():
    x = np.array([np.nan, np.inf, 1.0, 0.0, -0.0, -1.0, -np.inf, 5e-324, -5e-324, 1.7e308,
                  -1.7e308], dtype=np.float64)
    neg_nan = np.array([0xFFF8000000000000], dtype=np.uint64).view(np.float64)
    x = np.concatenate([x, neg_nan])
    bits = oracle.radix_bits(x, "f64")
    np.testing.assert_array_equal(bits, np_radix_bits(x, "f64"))
    back = oracle.radix_bits(bits, "f64", inverse=True)
    np.testing.assert_array_equal(back.view(np.uint64), x.view(np.uint64))
    order = x[np.argsort(bits, kind="stable")]
    assert np.isnan(order[0]) and np.signbit(order[0]) and np.isnan(order[-1])
    finite = order[1:-1]
    assert np.all(finite[:-1] <= finite[1:])
    iz = [i for i, v in enumerate(order) if v == 0]
    assert np.signbit(order[iz[0]]) and not np.signbit(order[iz[1]])


@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("greater", [False, True])
def test_merge_vs_numpy(kt, greater):
    na, nb = 7000, 9000
    a = oracle.stable_sort(keys_for(kt, na, 3)[np.arange(na) % 300], kt=kt, greater=greater)
    b = oracle.stable_sort(keys_for(kt, nb, 4)[np.arange(nb) % 500], kt=kt, greater=greater)
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    c, cv = oracle.merge(a, b, av, bv, kt, greater)
    cat = np.concatenate([a, b])
    order = np_stable_order(cat, kt, greater)
    np.testing.assert_array_equal(c, cat[order])
    np.testing.assert_array_equal(cv, np.concatenate([av, bv])[order])


def test_scans_u32_vs_numpy():
    x = oracle.generate(100000, "u32", 5)
    with np.errstate(over="ignore"):
        inc = np.cumsum(x.astype(np.uint64)) & 0xFFFFFFFF
    np.testing.assert_array_equal(oracle.inclusive_scan_u32(x, 7), (inc + 7) & 0xFFFFFFFF)
    exc = np.concatenate([np.zeros(1, np.uint64), inc[:-1]])
    np.testing.assert_array_equal(oracle.exclusive_scan_u32(x, 7), (exc + 7) & 0xFFFFFFFF)
    assert oracle.reduce_u32(x, 7) == (int(inc[-1]) + 7) & 0xFFFFFFFF


def test_int32_wrap_semantics():
    # Two's-complement wrap, computed as uint32 (the full-size int32 scan of
    # values in [0, 16) passes 2^31).
    x = np.full(1000, 2**30, dtype=np.int32)
    out = oracle.exclusive_scan_u32(x).view(np.int32)
    assert out[2] == -2**31 and out[4] == 0


def test_f64_scan_is_exact_for_generator():
    # Inputs are multiples of 2^-24 and prefixes stay below 2^29, so float64
    # prefix sums are exact: compare with integer arithmetic.
    x = oracle.generate(200000, "f32_01", 9)
    ints = (x.astype(np.float64) * 2 ** 24).astype(np.int64)
    assert np.all(ints.astype(np.float64) * 2.0 ** -24 == x)
    exact = np.cumsum(ints)
    np.testing.assert_array_equal(oracle.inclusive_scan_f64(x), exact.astype(np.float64) * 2.0 ** -24)
    np.testing.assert_array_equal(oracle.exclusive_scan_f64(x)[1:], oracle.inclusive_scan_f64(x)[:-1])
    assert oracle.reduce_f64(x) == float(exact[-1]) * 2.0 ** -24


def test_vk_scan_known_answer():
    # examples/vk_scan/vk_scan.cxx:22-23: host[i] = i, count 10000, exclusive.
    i = np.arange(10000, dtype=np.uint32)
    exp = (i.astype(np.uint64) * (i.astype(np.uint64) - 1) // 2).astype(np.uint32)
    exp[0] = 0
    np.testing.assert_array_equal(oracle.exclusive_scan_u32(i), exp)
    np.testing.assert_array_equal(oracle.scan_schedule_u32(i, 896), exp)


def np_segmented_scan(x, segs, inclusive, init, dtype):
    """Segmented scan by differences of one global cumsum (an algorithm
    unlike the oracle's per-segment std::scan)."""
    n = len(x)
    heads = np.zeros(n, dtype=bool)
    heads[0] = True
    s = np.asarray(segs, dtype=np.int64)
    heads[s[(s >= 0) & (s < n)]] = True
    seg_id = np.cumsum(heads) - 1
    start = np.flatnonzero(heads)
    c = np.concatenate([np.zeros(1, dtype), np.cumsum(x.astype(dtype), dtype=dtype)])
    first = start[seg_id]
    upto = np.arange(n) + (1 if inclusive else 0)
    return c[upto] - c[first] + dtype(init)


@pytest.mark.parametrize("inclusive", [False, True])
def test_segmented_scan_vs_numpy(inclusive):
    n = 50000
    x = oracle.generate(n, "u32", 8)
    lens = workloads.lognormal_lengths(400, 64.0, 1.5, 0, 4000, seed=3)
    segs = workloads.segment_starts(lens)
    segs = segs[segs <= n]
    segs = np.concatenate([segs, [n, n, n + 5]]).astype(np.int32)  # empty trailing segments
    with np.errstate(over="ignore"):
        exp = np_segmented_scan(x, segs, inclusive, 11, np.uint64) & np.uint64(0xFFFFFFFF)
    np.testing.assert_array_equal(oracle.segmented_scan_u32(x, segs, inclusive, 11), exp)
    f = oracle.generate(n, "f32_01", 9)
    ints = (f.astype(np.float64) * 2 ** 24).astype(np.int64)
    expf = np_segmented_scan(ints, segs, inclusive, 0, np.int64).astype(np.float64) * 2.0 ** -24
    np.testing.assert_array_equal(oracle.segmented_scan_f64(f, segs, inclusive, 0.0), expf)


def test_segmented_scan_special_cases():
    x = oracle.generate(3000, "i32_0_16", 2)
    # no segments, or one at 0: the plain scan
    for segs in (np.zeros(0, np.int32), np.zeros(1, np.int32)):
        np.testing.assert_array_equal(oracle.segmented_scan_u32(x, segs, False, 5),
                                      oracle.exclusive_scan_u32(x, 5))
        np.testing.assert_array_equal(oracle.segmented_scan_u32(x, segs, True, 5),
                                      oracle.inclusive_scan_u32(x, 5))
    # every element a segment: exclusive = init, inclusive = init + x
    every = np.arange(3000, dtype=np.int32)
    assert np.all(oracle.segmented_scan_u32(x, every, False, 5) == 5)
    np.testing.assert_array_equal(oracle.segmented_scan_u32(x, every, True, 5),
                                  x.view(np.uint32) + np.uint32(5))
    # duplicate starts (empty segments) change nothing
    segs = np.array([0, 100, 100, 100, 2000], dtype=np.int32)
    np.testing.assert_array_equal(oracle.segmented_scan_u32(x, segs, False, 0),
                                  oracle.segmented_scan_u32(x, np.array([0, 100, 2000], np.int32), False, 0))
    # a segment restarts the sum: element 100 of an exclusive scan is init
    assert oracle.segmented_scan_u32(x, segs, False, 3)[100] == 3


# --------------------------------------------------------- restatements

def brute_merge_path(a, b, diag):
    """A-count among the first diag outputs of the stable merge (ties to A)."""
    tags = np.concatenate([np.zeros(len(a), np.int8), np.ones(len(b), np.int8)])
    order = np.argsort(np.concatenate([a, b]), kind="stable")
    return int(np.sum(tags[order[:diag]] == 0))


def test_merge_path_brute_force():
    rng = np.random.default_rng(0)
    for trial in range(300):
        na, nb = rng.integers(0, 40, size=2)
        a = np.sort(rng.integers(0, 6, na)).astype(np.uint32)
        b = np.sort(rng.integers(0, 6, nb)).astype(np.uint32)
        for diag in range(na + nb + 1):
            assert oracle.merge_path(a, b, diag) == brute_merge_path(a, b, diag)


def brute_merge_path_upper(a, b, diag):
    """A-count among the first diag outputs of the merge whose ties take B
    first: the stable merge of (b, a), counting the a elements."""
    tags = np.concatenate([np.ones(len(b), np.int8), np.zeros(len(a), np.int8)])
    keys = np.concatenate([b, a])
    order = np.argsort(keys, kind="stable")
    return int((tags[order][:diag] == 0).sum())


def test_merge_path_upper_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(300):
        na, nb = rng.integers(0, 40, size=2)
        a = np.sort(rng.integers(0, 6, na)).astype(np.uint32)
        b = np.sort(rng.integers(0, 6, nb)).astype(np.uint32)
        for diag in range(na + nb + 1):
            assert oracle.merge_path(a, b, diag, upper=True) == brute_merge_path_upper(a, b, diag)


def test_merge_path_planted_upper_bound_is_caught():
    a = np.array([1, 2, 2, 3], np.uint32)
    b = np.array([2, 2, 4], np.uint32)
    diffs = [d for d in range(8) if oracle.merge_path(a, b, d, upper=True) != brute_merge_path(a, b, d)]
    assert diffs, "upper-bound merge path must disagree with the stable merge on ties"


def test_mergesort_range_frames():
    # Partitions of one coop group cover the group's two runs; the last one
    # ends at the runs' end (cta_mergesort.hxx:54-62).
    spacing, count = 7, 100
    for coop in (2, 4, 8, 16):
        for part in range(0, 16):
            r = oracle.mergesort_range(count, part, coop, spacing)
            start = (part // coop) * coop
            assert r[0] == min(count, spacing * start)
            assert r[1] == min(count, spacing * (start + coop // 2))
            assert r[2] == min(count, spacing * (start + coop // 2))
            assert r[3] == min(count, spacing * (start + coop))
    # With merge-path splits: consecutive partitions tile the output exactly.
    keys = np.sort(oracle.generate(100, "u32") % 9)
    keys = np.concatenate([np.sort(keys[:56]), np.sort(keys[56:])]).astype(np.uint32)
    coop, n_part = 16, 15
    mps = []
    for i in range(n_part + 1):
        r = oracle.mergesort_range(count, i, coop, spacing)
        diag = min(spacing * i, count) - r[0]
        mps.append(oracle.merge_path(keys[r[0]:r[1]], keys[r[2]:r[3]], diag))
    total = 0
    for cta in range(n_part):
        r = oracle.mergesort_range(count, cta, coop, spacing, mps[cta], mps[cta + 1])
        total += (r[1] - r[0]) + (r[3] - r[2])
    assert total == count


SCHED_SHAPES = [(0, 4, 3), (1, 4, 3), (11, 4, 3), (12, 4, 3), (13, 4, 3), (100, 4, 3),
                (1000, 8, 5), (4097, 16, 7), (3 * 896 + 5, 128, 7), (20000, 32, 15)]


@pytest.mark.parametrize("n,nt,vt", SCHED_SHAPES)
@pytest.mark.parametrize("kt", ["u32", "f32"])
def test_mergesort_schedule_equals_stable_sort(n, nt, vt, kt):
    keys = keys_for(kt, n, 6)
    keys = keys[np.arange(n) % 37] if n else keys
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.mergesort_schedule(keys, vals, nt, vt, kt)
    ek, ev = oracle.stable_sort(keys, vals, kt)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
    kg, vg = oracle.mergesort_schedule(keys, vals, nt, vt, kt, greater=True)
    ek, ev = oracle.stable_sort(keys, vals, kt, greater=True)
    np.testing.assert_array_equal(kg, ek)
    np.testing.assert_array_equal(vg, ev)


def test_mergesort_schedule_indices():
    keys = oracle.generate(5000, "u32", 7) % 100
    k, idx = oracle.mergesort_schedule(keys, None, 16, 7, sort_indices=True)
    np.testing.assert_array_equal(keys[idx], k)
    np.testing.assert_array_equal(idx, np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("plant", [1, 2])
def test_mergesort_schedule_planted_mistakes_are_caught(plant):
    n = 4097
    keys = oracle.generate(n, "u32", 8) % 50
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.stable_sort(keys, vals)
    with oracle.planted(plant):
        k, v = oracle.mergesort_schedule(keys, vals, 16, 7)
    assert not (np.array_equal(k, ek) and np.array_equal(v, ev))


@pytest.mark.parametrize("kt", KTS)
def test_radix_schedule(kt):
    n, nv = 30000, 4096
    keys = keys_for(kt, n, 10, low_entropy=True)
    vals = np.arange(n, dtype=np.uint32)
    (k, v), counts, offs = oracle.radix_schedule(keys, vals, nv, kt, want_first_pass=True)
    ek, ev = oracle.stable_sort_radix(keys, vals, kt)
    np.testing.assert_array_equal(k.view(np.uint8), ek.view(np.uint8))
    np.testing.assert_array_equal(v, ev)
    digit = (np_radix_bits(keys, kt) & 0xFF).astype(np.int64)
    tiles = (n + nv - 1) // nv
    bc = np.stack([np.bincount(digit[t * nv:(t + 1) * nv], minlength=256) for t in range(tiles)])
    np.testing.assert_array_equal(counts, bc)
    # Brute-force (bin, tile) offsets: keys of smaller digits, plus keys of
    # the same digit in earlier tiles.
    for t in range(tiles):
        for d in (0, 1, 17, 255):
            exp = int(np.sum(digit < d)) + int(np.sum(digit[:t * nv] == d))
            assert offs[t, d] == exp


def test_radix_schedule_planted_mistake_is_caught():
    n = 20000
    keys = oracle.generate(n, "u32", 11)
    with oracle.planted(3):
        k = oracle.radix_schedule(keys, None, 4096)
    assert not np.array_equal(k, oracle.stable_sort_radix(keys))


@pytest.mark.parametrize("n", [0, 1, 895, 896, 897, 8 * 896, 8 * 896 + 1, 100000])
def test_scan_schedule(n):
    x = oracle.generate(n, "i32_0_16", 12)
    np.testing.assert_array_equal(oracle.scan_schedule_u32(x, 896, 5), oracle.exclusive_scan_u32(x, 5))


def test_scan_schedule_planted_mistake_is_caught():
    x = oracle.generate(20000, "i32_0_16", 13)
    with oracle.planted(4):
        bad = oracle.scan_schedule_u32(x, 896)
    assert not np.array_equal(bad, oracle.exclusive_scan_u32(x))


def test_vk_radix2_known_answer_input():
    # examples/vk_radix2/vk_radix2.cxx:14-32: 8,388,608 rand() keys, std::sort.
    n = 256 * 32 * 32 * 32
    keys = oracle.glibc_rand(n)
    s = oracle.std_sort_u32(keys)
    np.testing.assert_array_equal(s, np.sort(keys))


@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("greater", [False, True])
def test_segmented_stable_sort_vs_numpy_lexsort(kt, greater):
    rng = np.random.default_rng(5)
    lens = rng.integers(0, 300, 200)
    lens[::17] = 0
    lens[::13] = 1
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int32)
    n = int(lens.sum())
    keys = keys_for(kt, n, 7)[np.arange(n) % 50]
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.segmented_stable_sort(keys, starts, vals, kt, greater)
    seg_id = np.repeat(np.arange(len(lens)), lens)
    if greater:
        key_rank = ~keys if kt in ("u32", "u64") else -keys.astype(np.float64)
    else:
        key_rank = keys
    order = np.lexsort((key_rank, seg_id))
    np.testing.assert_array_equal(v, vals[order])
    np.testing.assert_array_equal(k, keys[order])


def test_segmented_stable_sort_starts_past_the_end():
    """Starts at or beyond n give empty segments: the result equals the sort
    over the starts below n (numpy lexsort by segment id, then key), and a
    start far past n reads and writes nothing outside the arrays."""
    n = 5000
    keys = oracle.generate(n, "u32", 21)[np.arange(n) % 37]
    vals = np.arange(n, dtype=np.uint32)
    starts = np.array([0, 7, 7, 1000, n // 2, n, n, n + 9, 2 ** 31 - 1], dtype=np.int32)
    k, v = oracle.segmented_stable_sort(keys, starts, vals, "u32")
    inside = starts[starts < n]
    seg_id = np.searchsorted(inside, np.arange(n), side="right")
    order = np.lexsort((keys, seg_id))
    np.testing.assert_array_equal(v, vals[order])
    np.testing.assert_array_equal(k, keys[order])
    k2, v2 = oracle.segmented_stable_sort(keys, inside, vals, "u32")
    np.testing.assert_array_equal(k, k2)
    np.testing.assert_array_equal(v, v2)


def test_oracle_rejects_values_of_the_wrong_length():
    """The oracle passes raw buffers to C: a value array shorter than its
    keys must be an error, not a read past its end."""
    a = np.arange(5, dtype=np.uint32)
    with pytest.raises(ValueError):
        oracle.merge(a, a, np.arange(5, dtype=np.uint32), np.arange(3, dtype=np.uint32))
    with pytest.raises(ValueError):
        oracle.stable_sort(a, np.arange(4, dtype=np.uint32))
    with pytest.raises(ValueError):
        oracle.stable_sort_radix(a, np.arange(6, dtype=np.uint32))


def test_oracle_under_address_and_undefined_behavior_sanitizers():
    """Every exported oracle function, on empty, tiny, ragged and tie-heavy
    inputs of every key type, built with -fsanitize=address,undefined
    (oracle/sanitize_main.cpp): an out-of-bounds access or undefined
    operation in the checker fails here, and the driver's own cross-checks
    (schedules vs std:: algorithms, merge path vs the merge's sources) must
    all pass."""
    import os
    import shutil
    import subprocess
    if not shutil.which("g++") or not shutil.which("make"):
        pytest.skip("g++ and make are needed to build the sanitizer driver")
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    r = subprocess.run(["make", "-s", "-C", root, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "oracle sanitize: ALL PASS" in r.stdout


@pytest.mark.parametrize("kt", ["u32", "f64"])
@pytest.mark.parametrize("greater", [False, True])
@pytest.mark.parametrize("n,threads", [(0, 4), (1, 4), (7, 3), (5000, 8), (100003, 5),
                                        (65536, 16)])
def test_stable_sort_threads_equals_stable_sort(kt, greater, n, threads):
    """The multi-threaded CPU baseline (chunk sorts + merge-path-split merge
    rounds) gives the stable sort: checked against numpy's stable argsort on
    tie-heavy keys, and against the single-threaded oracle."""
    keys = keys_for(kt, n, 9, low_entropy=True)
    keys = keys[np.arange(n) % 61] if n else keys
    vals = np.arange(n, dtype=np.uint32)
    (k, v), dt = oracle.stable_sort_threads(keys, vals, kt, greater, threads)
    assert dt >= 0.0
    order = np_stable_order(keys, kt, greater)
    np.testing.assert_array_equal(v, vals[order])
    np.testing.assert_array_equal(k, keys[order])
    ek, ev = oracle.stable_sort(keys, vals, kt, greater)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
