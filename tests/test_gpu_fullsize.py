"""GPU parity at the full workload sizes through size-independent properties:
stable-sort characterisation (sorted + permutation + gather + tie order),
merge against an independent device sort, scans against exact integer /
float64 prefix sums computed by torch (exact for these generators)."""
import pytest
import torch

import mgpu_hip
from gpu_util import as_signed_order, banner, check_stable_sort_properties, gen_dev

pytestmark = pytest.mark.gpu

N_SORT = 1 << 28


@pytest.mark.parametrize("kind", ["u32", "lowentropy"])
@pytest.mark.parametrize("algo", ["radix_count_scan_scatter", "radix_single_pass", "mergesort",
                                  "mergesort_two_way"])
def test_sort_2p28_kv(kind, algo):
    banner("fullsize sort", algo, kind)
    keys_in = gen_dev(N_SORT, kind, 0)
    keys = keys_in.clone()
    vals = gen_dev(N_SORT, "iota")
    if algo.startswith("radix"):
        old = mgpu_hip.radix_algorithm()
        mgpu_hip.set_radix_algorithm(mgpu_hip.RADIX_SINGLE_PASS if algo == "radix_single_pass"
                                     else mgpu_hip.RADIX_COUNT_SCAN_SCATTER)
        try:
            mgpu_hip.radix_sort(keys, vals)
        finally:
            mgpu_hip.set_radix_algorithm(old)
    elif algo == "mergesort_two_way":
        old = mgpu_hip.mergesort_passes()
        mgpu_hip.set_mergesort_passes(mgpu_hip.MERGESORT_TWO_WAY)
        try:
            mgpu_hip.mergesort_kv(keys, vals)
        finally:
            mgpu_hip.set_mergesort_passes(old)
    else:
        mgpu_hip.mergesort_kv(keys, vals)  # automatic: four-way passes at this size
    torch.cuda.synchronize()
    check_stable_sort_properties(keys_in, keys, vals)


def test_merge_2x2p27():
    banner("fullsize merge")
    n = 1 << 27
    a = gen_dev(n, "u32", 1)
    b = gen_dev(n, "u32", 2)
    mgpu_hip.radix_sort(a)
    mgpu_hip.radix_sort(b)
    av = gen_dev(n, "iota")
    bv = gen_dev(n, "iota", 0, n)
    c = torch.empty(2 * n, dtype=torch.uint32, device="cuda")
    cv = torch.empty(2 * n, dtype=torch.uint32, device="cuda")
    mgpu_hip.merge(a, av, b, bv, c, cv, "less")
    torch.cuda.synchronize()
    check_stable_sort_properties(torch.cat([a, b]), c, cv)
    ref = torch.sort(as_signed_order(torch.cat([a, b])), stable=True).values
    assert bool((ref == as_signed_order(c)).all())


def test_scan_i32_2p29():
    banner("fullsize scan i32")
    n = 1 << 29
    x = gen_dev(n, "i32_0_16", 0)
    out = torch.empty_like(x)
    mgpu_hip.scan_into(x, out)
    exact = torch.cumsum(x.long(), 0)
    exact = torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), exact[:-1]])
    assert bool(((exact & 0xFFFFFFFF) == (out.long() & 0xFFFFFFFF)).all())
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    mgpu_hip.reduce(x, total)
    assert int(total.long().item() & 0xFFFFFFFF) == int(x.long().sum().item()) & 0xFFFFFFFF


def test_scan_f32_2p29():
    banner("fullsize scan f32")
    n = 1 << 29
    x = gen_dev(n, "f32_01", 0)
    out = torch.empty_like(x)
    mgpu_hip.scan_into(x, out, inclusive=True)
    # Multiples of 2^-24 with prefix < 2^29: float64 cumsum is exact in any order.
    exact = torch.cumsum(x.double(), 0)
    nz = exact > 0
    rel = ((out.double() - exact).abs()[nz] / exact[nz]).max().item()
    print("f32 scan 2^29 max rel err", rel, flush=True)
    assert rel <= 2e-6
    total = torch.zeros(1, dtype=torch.float32, device="cuda")
    mgpu_hip.reduce(x, total)
    ex = exact[-1].item()
    assert abs(total.item() - ex) <= ex * 2 ** -24 * 1.01
