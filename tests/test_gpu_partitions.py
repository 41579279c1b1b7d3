"""GPU parity: merge_path_partitions<bounds> (inc/mgpu/vk/partition.hxx:9-20)
against the oracle's merge_path (inc/mgpu/common/cta_merge.hxx:10-31), which
the CPU tests pin by brute force for both bounds.  Tie-heavy sorted inputs,
both comparators, several spacings, empty inputs, and disjoint ranges (the
co-ranks at the grid's edges)."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host

pytestmark = pytest.mark.gpu

KTS = ("u32", "i32", "u64", "f32", "i64", "f64")


def sorted_keys(kt, n, seed, greater, mod=23):
    kind = {"u32": "u32", "i32": "u32", "u64": "u64", "f32": "f32_pm1", "i64": "u64",
            "f64": "u64"}[kt]
    x = oracle.wide_keys(kt, oracle.generate(n, kind, seed))
    x = np.sort(x[np.arange(n) % mod] if n else x)
    return x[::-1].copy() if greater else x


def expected(a, b, spacing, kt, greater, upper):
    total = len(a) + len(b)
    n = mgpu_hip.num_merge_partitions(total, spacing)
    return np.array([oracle.merge_path(a, b, min(i * spacing, total), kt, greater, upper)
                     for i in range(n)], dtype=np.int64)


@pytest.mark.parametrize("kt", KTS)
@pytest.mark.parametrize("bounds", ["lower", "upper"])
@pytest.mark.parametrize("comp", ["less", "greater"])
def test_partitions_vs_oracle(kt, bounds, comp):
    greater, upper = comp == "greater", bounds == "upper"
    for na, nb, spacing in ((3000, 5000, 97), (4099, 17, 512), (1, 1, 1), (700, 700, 1)):
        banner("merge_path_partitions", kt, bounds, comp, na, nb, spacing)
        a = sorted_keys(kt, na, 1, greater)
        b = sorted_keys(kt, nb, 2, greater)
        mp = mgpu_hip.merge_path_partitions(to_dev(a), to_dev(b), spacing, comp, bounds)
        np.testing.assert_array_equal(to_host(mp), expected(a, b, spacing, kt, greater, upper))


@pytest.mark.parametrize("bounds", ["lower", "upper"])
def test_partitions_edges(bounds):
    upper = bounds == "upper"
    banner("merge_path_partitions edges", bounds)
    # Empty inputs, and disjoint ranges: every A key below (then above) every
    # B key puts the co-ranks on the edges of the grid.
    cases = [(np.zeros(0, np.uint32), np.arange(10, dtype=np.uint32)),
             (np.arange(10, dtype=np.uint32), np.zeros(0, np.uint32)),
             (np.arange(1000, dtype=np.uint32), np.arange(1000, 2000, dtype=np.uint32)),
             (np.arange(1000, 2000, dtype=np.uint32), np.arange(1000, dtype=np.uint32)),
             (np.full(900, 5, np.uint32), np.full(800, 5, np.uint32))]
    for a, b in cases:
        mp = mgpu_hip.merge_path_partitions(to_dev(a), to_dev(b), 64, "less", bounds)
        np.testing.assert_array_equal(to_host(mp), expected(a, b, 64, "u32", False, upper))
    # All keys equal: lower takes every A first, upper every B first.
    a, b = np.full(900, 5, np.uint32), np.full(800, 5, np.uint32)
    mp = to_host(mgpu_hip.merge_path_partitions(to_dev(a), to_dev(b), 100, "less", bounds))
    diag = np.minimum(np.arange(len(mp)) * 100, 1700)
    exp = np.maximum(diag - 800, 0) if upper else np.minimum(diag, 900)
    np.testing.assert_array_equal(mp, exp)


def test_partitions_argument_errors():
    banner("merge_path_partitions errors")
    lib = mgpu_hip.load()
    a = torch.zeros(8, dtype=torch.uint32, device="cuda")
    mp = torch.zeros(16, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    f = lib.mgpu_merge_path_partitions_u32
    assert f(s, a.data_ptr(), 8, a.data_ptr(), 8, mp.data_ptr(), 0, 0, 0) == mgpu_hip.MGPU_INVALID_ARG
    assert f(s, a.data_ptr(), -1, a.data_ptr(), 8, mp.data_ptr(), 4, 0, 0) == mgpu_hip.MGPU_INVALID_ARG
    assert f(s, a.data_ptr(), 8, a.data_ptr(), 8, mp.data_ptr(), 4, 0, 7) == mgpu_hip.MGPU_INVALID_ARG
    assert f(s, a.data_ptr(), 8, a.data_ptr(), 8, None, 4, 0, 0) == mgpu_hip.MGPU_INVALID_ARG
    torch.cuda.synchronize()
    assert bool((mp == 0).all())
