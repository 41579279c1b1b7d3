"""GPU parity above 2^31 elements, where 32-bit index arithmetic would wrap,
through size-independent properties:
* radix sort of 2^31 + 12345 and of 2^32 - 12289 (its count limit) uint32
  keys: ordered, and the same multiset as
  the input (sum, and sums of products of the 16-bit halves);
* exclusive scan of 2^31 + 7 int32 in [0, 16): out[0] = init and
  out[i + 1] = out[i] + x[i] (mod 2^32) for every i, which fixes every output;
* merge of two sorted 2^30-key arrays: ordered, and the multiset of both;
* key-value mergesort, segmented scan and key-value segmented sort at
  2^31 - 1 elements, their largest accepted count."""
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


def test_radix_keys_at_count_limit():
    """2^32 - 12289 keys, the largest count the 32-bit radix entry points
    accept (output positions are 32-bit, less one 12288-key tile): ordered,
    and the same multiset as the input."""
    n = (1 << 32) - 12289
    banner("radix keys at the count limit", n)
    keys = gen_dev(n, "u32", 46)
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


def test_mergesort_kv_at_count_limit():
    """Key-value mergesort of 2^31 - 1 pairs (the largest count the entry
    point accepts), keys in [0, 2^20) so that nearly every key has ties and
    values = input index.  Checked in chunks: keys ordered; values a
    permutation of [0, n) (every slot of a bitmap hit); keys_out[i] ==
    keys_in[vals_out[i]]; equal keys keep ascending values.  Together these
    fix the stable sort uniquely."""
    n = (1 << 31) - 1
    banner("mergesort kv at the count limit", n)
    keys = gen_dev(n, "u32", 45)
    keys.view(torch.int32).bitwise_and_(0xFFFFF)
    keys_in = keys.clone()
    vals = gen_dev(n, "iota", 0)
    mgpu_hip.mergesort_kv(keys, vals)
    torch.cuda.synchronize()
    assert_ordered(keys)
    seen = torch.zeros(n, dtype=torch.uint8, device=keys.device)
    kin = keys_in.view(torch.int32)
    for i in range(0, n, CHUNK):
        j = min(i + CHUNK, n)
        v = vals[i:j].view(torch.int32).long()
        seen[v] = 1
        assert bool((kin[v] == keys.view(torch.int32)[i:j]).all()), \
            "keys_out != keys_in[vals_out] near %d" % i
        del v
    assert bool(seen.all()), "values are not a permutation"
    del seen, keys_in, kin
    for i in range(0, n - 1, CHUNK):
        j = min(i + CHUNK + 1, n)
        k = keys.view(torch.int32)[i:j]
        v = vals.view(torch.int32)[i:j]
        tie = k[1:] == k[:-1]
        assert bool((v[1:][tie] > v[:-1][tie]).all()), "ties out of input order near %d" % i
        del k, v, tie
    del keys, vals
    torch.cuda.empty_cache()


def random_starts(n, count, seed):
    """Ascending distinct segment starts in [0, n), element 0 always one."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    s = torch.unique(torch.randint(1, n, (count - 1,), generator=g, device="cuda"))
    return torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), s]).to(torch.int32)


def start_mask(n, segs):
    m = torch.zeros(n, dtype=torch.bool, device="cuda")
    m[segs.long()] = True
    return m


def test_segmented_scan_at_count_limit():
    """Exclusive segmented scan of 2^31 - 1 int32 in [0, 16) over ~65536
    random segments, init 3: out[s] = 3 at every segment start s and
    out[p] = out[p - 1] + x[p - 1] everywhere else, which fixes every output."""
    n = (1 << 31) - 1
    banner("segmented scan at the count limit", n)
    x = gen_dev(n, "i32_0_16", 47)
    segs = random_starts(n, 1 << 16, 47)
    out = torch.empty_like(x)
    mgpu_hip.segmented_scan(x, out, segs, inclusive=False, init=3)
    torch.cuda.synchronize()
    start = start_mask(n, segs)
    assert bool((out[segs.long()] == 3).all()), "segment start != init"
    for i in range(1, n, CHUNK):
        j = min(i + CHUNK, n)
        d = out[i:j] - out[i - 1:j - 1] - x[i - 1:j - 1]
        assert bool((d[~start[i:j]] == 0).all()), "out[p] != out[p-1] + x[p-1] near %d" % i
        del d
    del x, out, start
    torch.cuda.empty_cache()


def test_segmented_sort_kv_at_count_limit():
    """Key-value segmented sort of 2^31 - 1 pairs over ~2^18 random segments
    (lengths from 1 to a few times 8192, so every size bucket and the
    large-segment radix path take part), keys in [0, 2^16) for ties, values =
    input index.  Checked in chunks: values a permutation of [0, n), each in
    the segment of its output position; keys_out[i] == keys_in[vals_out[i]];
    inside a segment keys ascend and equal keys keep ascending values.
    Together these fix every segment's stable sort."""
    n = (1 << 31) - 1
    banner("segmented sort kv at the count limit", n)
    keys = gen_dev(n, "u32", 48)
    keys.view(torch.int32).bitwise_and_(0xFFFF)
    keys_in = keys.clone()
    vals = gen_dev(n, "iota", 0)
    segs = random_starts(n, 1 << 18, 48)
    mgpu_hip.segmented_sort(keys, segs, vals)
    torch.cuda.synchronize()
    ki, ko, vo = keys_in.view(torch.int32), keys.view(torch.int32), vals.view(torch.int32)
    sl = segs.long()
    seen = torch.zeros(n, dtype=torch.uint8, device="cuda")
    for i in range(0, n, CHUNK):
        j = min(i + CHUNK, n)
        v = vo[i:j].long()
        seen[v] = 1
        assert bool((ki[v] == ko[i:j]).all()), "keys_out != keys_in[vals_out] near %d" % i
        pos = torch.arange(i, j, device="cuda")
        assert bool((torch.searchsorted(sl, v, right=True) ==
                     torch.searchsorted(sl, pos, right=True)).all()), \
            "a value left its segment near %d" % i
        del v, pos
    assert bool(seen.all()), "values are not a permutation"
    del seen, keys_in, ki
    start = start_mask(n, segs)
    for i in range(0, n - 1, CHUNK):
        j = min(i + CHUNK + 1, n)
        k, v = ko[i:j], vo[i:j]
        ok = (k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (v[1:] > v[:-1]))
        assert bool(ok[~start[i + 1:j]].all()), "segment not stably sorted near %d" % i
        del k, v, ok
    del keys, vals, start
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
