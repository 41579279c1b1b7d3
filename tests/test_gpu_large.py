"""GPU parity above 2^31 elements, where 32-bit index arithmetic would wrap,
through size-independent properties:
* radix sort of 2^31 + 12345 uint32 keys: ordered, and the same multiset as
  the input (sum, and sums of products of the 16-bit halves);
* exclusive scan of 2^31 + 7 int32 in [0, 16): out[0] = init and
  out[i + 1] = out[i] + x[i] (mod 2^32) for every i, which fixes every output;
* merge of two sorted 2^30-key arrays: ordered, and the multiset of both."""
import pytest
import torch

import mgpu_hip
from gpu_util import as_signed_order, banner, gen_dev

pytestmark = pytest.mark.gpu

CHUNK = 1 << 28


def multiset_signature(t):
    """(sum x, sum lo*hi, sum lo^2, sum hi^2) over uint32 x = hi * 2^16 + lo,
    accumulated exactly in chunks (each chunk's sums fit int64)."""
    sig = [0, 0, 0, 0]
    for i in range(0, t.numel(), CHUNK):
        x = t[i:i + CHUNK].view(torch.int32).long() & 0xFFFFFFFF
        lo, hi = x & 0xFFFF, x >> 16
        sig[0] += int(x.sum())
        sig[1] += int((lo * hi).sum())
        sig[2] += int((lo * lo).sum())
        sig[3] += int((hi * hi).sum())
        del x, lo, hi
    return sig


def assert_ordered(t):
    for i in range(0, t.numel() - 1, CHUNK):
        s = as_signed_order(t[i:i + CHUNK + 1])
        assert bool((s[1:] >= s[:-1]).all()), "not ordered near %d" % i
        del s


def test_radix_keys_above_2p31():
    n = (1 << 31) + 12345
    banner("radix keys", n)
    keys = gen_dev(n, "u32", 41)
    before = multiset_signature(keys)
    mgpu_hip.radix_sort(keys)
    torch.cuda.synchronize()
    assert_ordered(keys)
    assert multiset_signature(keys) == before
    del keys
    torch.cuda.empty_cache()


def test_exclusive_scan_above_2p31():
    n = (1 << 31) + 7
    banner("scan i32", n)
    x = gen_dev(n, "i32_0_16", 42)
    out = torch.empty_like(x)
    mgpu_hip.scan_into(x, out, inclusive=False, init=-5)
    torch.cuda.synchronize()
    assert int(out[0]) == -5
    for i in range(0, n - 1, CHUNK):
        j = min(i + CHUNK, n - 1)
        # int32 arithmetic wraps modulo 2^32 on the device, like the scan.
        d = out[i + 1:j + 1] - out[i:j] - x[i:j]
        assert bool((d == 0).all()), "out[i+1] != out[i] + x[i] near %d" % i
        del d
    del x, out
    torch.cuda.empty_cache()


def test_merge_2x2p30():
    m = 1 << 30
    banner("merge", 2 * m)
    a = gen_dev(m, "u32", 43)
    b = gen_dev(m, "u32", 44)
    mgpu_hip.radix_sort(a)
    mgpu_hip.radix_sort(b)
    c = torch.empty(2 * m, dtype=torch.uint32, device="cuda")
    mgpu_hip.merge_keys(a, b, c, "less")
    torch.cuda.synchronize()
    assert_ordered(c)
    sa, sb = multiset_signature(a), multiset_signature(b)
    assert multiset_signature(c) == [p + q for p, q in zip(sa, sb)]
    del a, b, c
    torch.cuda.empty_cache()
