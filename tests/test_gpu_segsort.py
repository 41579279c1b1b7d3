"""GPU parity: batched segmented sort against one std::stable_sort per segment
(bit-exact), across bucket boundaries, key types, comparators, the tuned
table, and graph replays."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host
from mgpu_hip import workloads

pytestmark = pytest.mark.gpu

KTS = ("u32", "i32", "u64", "f32", "i64", "f64")


def make_keys(kt, n, seed, ties=True):
    kind = {"u32": "u32", "i32": "u32", "u64": "u64", "f32": "f32_pm1", "i64": "u64",
            "f64": "u64"}[kt]
    x = oracle.wide_keys(kt, oracle.generate(n, kind, seed))
    if ties and n:
        x = x[np.arange(n) % 211]
    return x.copy()


def boundary_lengths():
    """Lengths on both sides of every bucket capacity of the default tables
    (128 ... 8192, and 16384 where the 1024 x 17 tile fits) and of every
    blocksort shape capacity, and of one to four large-segment radix tiles
    (8192)."""
    caps = [128, 192, 256, 320, 384, 448, 512, 576, 768, 896, 1024, 1152, 1536, 1664, 2048,
            2304, 3072, 3328, 4096, 4352, 6144, 6656, 8192, 8704, 12288, 16384, 17408]
    lens = [0, 1, 2, 3, 0, 1]
    for c in caps:
        lens += [c - 1, c, c + 1]
    lens += [2 * 8192 + 5, 3 * 8192, 3 * 12288, 70000]
    return np.array(lens, dtype=np.int64)


def run_case(lens, kt, comp, kv, seed, tuning=None):
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    keys = make_keys(kt, n, seed)
    vals = np.arange(n, dtype=np.uint32)
    dk = to_dev(keys)
    dv = to_dev(vals) if kv else None
    mgpu_hip.segmented_sort(dk, to_dev(starts), dv, comp=comp, tuning=tuning)
    greater = comp == "greater"
    if kv:
        ek, ev = oracle.segmented_stable_sort(keys, starts, vals, kt, greater)
        np.testing.assert_array_equal(to_host(dv), ev)
    else:
        ek = oracle.segmented_stable_sort(keys, starts, None, kt, greater)
    np.testing.assert_array_equal(to_host(dk).view(np.uint8), ek.view(np.uint8))


@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("comp", ["less", "greater"])
@pytest.mark.parametrize("kv", [True, False])
def test_segsort_bucket_boundaries(kt, comp, kv):
    banner("segsort boundaries", kt, comp, kv)
    run_case(boundary_lengths(), kt, comp, kv, 1)


@pytest.mark.parametrize("seed", [0, 1])
def test_segsort_lognormal_u32_kv(seed):
    banner("segsort lognormal", seed)
    lens = workloads.lognormal_lengths(4096, 256.0, 1.5, 1, 1 << 17, seed=seed)
    run_case(lens, "u32", "less", True, 10 + seed)


def test_segsort_single_large_and_tiny_segments():
    banner("segsort single large")
    run_case(np.array([1 << 20], dtype=np.int64), "u32", "less", True, 3)
    run_case(np.ones(5000, dtype=np.int64), "u32", "less", True, 4)
    run_case(np.full(3000, 2, dtype=np.int64), "f32", "greater", True, 5)


def test_segsort_float_signed_zero_ties_like_comparator():
    banner("segsort signed zero")
    # -0.0 and +0.0 compare equal, so a stable sort keeps their input order
    # in every bucket, including the radix path of large segments.
    lens = np.array([300, 9000, 20000], dtype=np.int64)
    n = int(lens.sum())
    keys = np.where(np.arange(n) % 2 == 0, np.float32(0.0), np.float32(-0.0)).astype(np.float32)
    keys[::7] = 1.0
    starts = workloads.segment_starts(lens)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = to_dev(keys), to_dev(vals)
    mgpu_hip.segmented_sort(dk, to_dev(starts), dv)
    ek, ev = oracle.segmented_stable_sort(keys, starts, vals, "f32")
    np.testing.assert_array_equal(to_host(dv), ev)
    np.testing.assert_array_equal(to_host(dk).view(np.uint32), ek.view(np.uint32))


def test_segsort_autotuned_table():
    banner("segsort autotune")
    t = mgpu_hip.segmented_sort_autotune("u32", True, "less")
    shapes = mgpu_hip.segsort_shapes()
    for cap, sh in zip(t["cap"], t["shape"]):
        assert shapes[sh][0] * shapes[sh][1] >= cap
    print("tuned", t["cap"], [shapes[s] for s in t["shape"]], flush=True)
    lens = workloads.lognormal_lengths(3000, 256.0, 1.5, 1, 1 << 16, seed=7)
    run_case(lens, "u32", "less", True, 20, tuning=t)
    run_case(boundary_lengths(), "u32", "less", True, 21)  # library table after tuning


def test_segsort_graph_replay():
    """The whole segmented sort captured once and replayed; each replay sorts
    freshly restored input (eager copies between replays)."""
    banner("segsort graph replay")
    lens = workloads.lognormal_lengths(2000, 256.0, 1.5, 1, 1 << 16, seed=9)
    lens[5] = 40000
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    stream = torch.cuda.Stream()
    segs = to_dev(starts)
    expect = {}
    src = {}
    for r in range(3):
        keys = make_keys("u32", n, 100 + r)
        src[r] = (to_dev(keys), to_dev(np.arange(n, dtype=np.uint32)))
        expect[r] = oracle.segmented_stable_sort(keys, starts, np.arange(n, dtype=np.uint32), "u32")
    dk = torch.empty(n, dtype=torch.uint32, device="cuda")
    dv = torch.empty(n, dtype=torch.uint32, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        dk.copy_(src[0][0])
        dv.copy_(src[0][1])
        aux = mgpu_hip.segmented_sort(dk, segs, dv, stream=stream)  # eager warm-up
        stream.synchronize()
        g = mgpu_hip.Graph(stream)
        with g:
            mgpu_hip.segmented_sort(dk, segs, dv, aux=aux, stream=stream)
    stream.synchronize()
    for r in range(3):
        with torch.cuda.stream(stream):
            dk.copy_(src[r][0])
            dv.copy_(src[r][1])
            g.launch()
        stream.synchronize()
        np.testing.assert_array_equal(to_host(dk), expect[r][0])
        np.testing.assert_array_equal(to_host(dv), expect[r][1])


@pytest.mark.parametrize("kv", [True, False])
def test_segsort_starts_past_the_end(kv):
    """Starts at or beyond count are empty segments (as in the segmented
    scan): the sort equals the oracle's over the same starts, a segment that
    reaches count goes through the large-segment radix path, and guard words
    around the arrays stay untouched."""
    n = 50000
    banner("segsort starts past the end", n, kv)
    nv = 512 * 17
    starts = np.array([0, 7, 7, 7, nv, nv, n // 2, n, n, n + 9], dtype=np.int32)
    keys = make_keys("u32", n, 31)
    vals = np.arange(n, dtype=np.uint32)
    guard = 64
    gk = torch.full((n + 2 * guard,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    gk[guard:guard + n] = to_dev(keys.view(np.int32))
    gv = torch.full((n + 2 * guard,), 0x3C3C3C3C, dtype=torch.int32, device="cuda")
    gv[guard:guard + n] = to_dev(vals.view(np.int32))
    dk = gk[guard:guard + n].view(torch.uint32)
    dv = gv[guard:guard + n].view(torch.uint32) if kv else None
    mgpu_hip.segmented_sort(dk, to_dev(starts), dv)
    hk, hv = to_host(gk), to_host(gv)
    assert np.all(hk[:guard] == 0x5A5A5A5A) and np.all(hk[guard + n:] == 0x5A5A5A5A)
    assert np.all(hv[:guard] == 0x3C3C3C3C) and np.all(hv[guard + n:] == 0x3C3C3C3C)
    if kv:
        ek, ev = oracle.segmented_stable_sort(keys, starts, vals, "u32")
        np.testing.assert_array_equal(hv[guard:guard + n].view(np.uint32), ev)
    else:
        ek = oracle.segmented_stable_sort(keys, starts, None, "u32")
    np.testing.assert_array_equal(hk[guard:guard + n].view(np.uint32), ek)
