"""GPU parity on device pointers that are not 16-byte aligned (views at an
element offset into a larger buffer), for every primitive with a vector
(128-bit) fast path or tile-aligned loads.  Sentinels on both sides of each
view must be untouched."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host
from mgpu_hip import workloads

pytestmark = pytest.mark.gpu

OFFSETS = [1, 2, 3]
PAD = 5


def padded(x, off):
    """Device buffer [off + PAD sentinels][x][PAD sentinels] and the view of x."""
    sentinel = np.full(off + PAD, 0x5A, dtype=x.dtype) if x.dtype.kind in "ui" else \
        np.full(off + PAD, -7.5, dtype=x.dtype)
    buf = to_dev(np.concatenate([sentinel, x, sentinel[:PAD]]))
    return buf, buf[off + PAD: off + PAD + len(x)], sentinel


def check_guards(buf, off, n, sentinel):
    h = to_host(buf)
    np.testing.assert_array_equal(h[: off + PAD], sentinel)
    np.testing.assert_array_equal(h[off + PAD + n:], sentinel[:PAD])


@pytest.mark.parametrize("off", OFFSETS)
def test_radix_kv_misaligned(off):
    banner("radix misaligned", off)
    n = 3 * 512 * 8 + 77
    keys = oracle.generate(n, "u32", 11)[np.arange(n) % 4099]
    vals = np.arange(n, dtype=np.uint32)
    kb, kv, ks = padded(keys, off)
    vb, vv, vs = padded(vals, (off + 1) % 4)
    mgpu_hip.radix_sort(kv, vv)
    ek, ev = oracle.stable_sort_radix(keys, vals, "u32")
    np.testing.assert_array_equal(to_host(kv), ek)
    np.testing.assert_array_equal(to_host(vv), ev)
    check_guards(kb, off, n, ks)
    check_guards(vb, (off + 1) % 4, n, vs)


@pytest.mark.parametrize("off", OFFSETS)
def test_mergesort_kv_misaligned(off):
    banner("mergesort misaligned", off)
    n = 2 * 512 * 19 + 1001
    keys = oracle.generate(n, "f32_pm1", 12)[np.arange(n) % 997]
    vals = np.arange(n, dtype=np.uint32)
    kb, kv, ks = padded(keys, off)
    vb, vv, vs = padded(vals, off)
    mgpu_hip.mergesort_kv(kv, vv)
    ek, ev = oracle.stable_sort(keys, vals, "f32")
    np.testing.assert_array_equal(to_host(kv).view(np.uint32), ek.view(np.uint32))
    np.testing.assert_array_equal(to_host(vv), ev)
    check_guards(kb, off, n, ks)
    check_guards(vb, off, n, vs)


@pytest.mark.parametrize("off", OFFSETS)
def test_merge_kv_misaligned(off):
    banner("merge misaligned", off)
    na, nb = 20011, 13007
    a = np.sort(oracle.generate(na, "u32", 13) % 5000)
    b = np.sort(oracle.generate(nb, "u32", 14) % 5000)
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    _, da, _ = padded(a, off)
    _, dav, _ = padded(av, (off + 2) % 4)
    _, db, _ = padded(b, (off + 1) % 4)
    _, dbv, _ = padded(bv, off)
    cb, dc, cs = padded(np.zeros(na + nb, np.uint32), (off + 3) % 4)
    cvb, dcv, cvs = padded(np.zeros(na + nb, np.uint32), off)
    mgpu_hip.merge(da, dav, db, dbv, dc, dcv, "less")
    ek, ev = oracle.merge(a, b, av, bv)
    np.testing.assert_array_equal(to_host(dc), ek)
    np.testing.assert_array_equal(to_host(dcv), ev)
    check_guards(cb, (off + 3) % 4, na + nb, cs)
    check_guards(cvb, off, na + nb, cvs)


@pytest.mark.parametrize("off", OFFSETS)
def test_segmented_sort_and_scan_misaligned(off):
    banner("segmented misaligned", off)
    lens = workloads.lognormal_lengths(500, 256.0, 1.5, 1, 30000, seed=off)
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    keys = oracle.generate(n, "u32", 15)[np.arange(n) % 211]
    vals = np.arange(n, dtype=np.uint32)
    kb, kv, ks = padded(keys, off)
    vb, vv, vs = padded(vals, off)
    mgpu_hip.segmented_sort(kv, to_dev(starts), vv)
    ek, ev = oracle.segmented_stable_sort(keys, starts, vals, "u32")
    np.testing.assert_array_equal(to_host(kv), ek)
    np.testing.assert_array_equal(to_host(vv), ev)
    check_guards(kb, off, n, ks)
    check_guards(vb, off, n, vs)
    x = oracle.generate(n, "f32_01", 16)
    xb, xv, xs = padded(x, off)
    yb, yv, ys = padded(np.zeros(n, np.float32), (off + 2) % 4)
    mgpu_hip.segmented_scan(xv, yv, to_dev(starts), inclusive=True)
    exact = oracle.segmented_scan_f64(x, starts, True, 0.0)
    got = to_host(yv).astype(np.float64)
    nz = exact > 0
    assert np.max(np.abs(got[nz] - exact[nz]) / exact[nz]) <= 2e-6
    check_guards(xb, off, n, xs)
    check_guards(yb, (off + 2) % 4, n, ys)


@pytest.mark.parametrize("off", OFFSETS)
def test_scan_out_of_place_misaligned_u32_and_u64(off):
    banner("scan misaligned out of place", off)
    n = 5 * 256 * 80 + 333
    x = oracle.generate(n, "u32", 17)
    xb, xv, xs = padded(x, off)
    yb, yv, ys = padded(np.zeros(n, np.uint32), (off + 1) % 4)
    mgpu_hip.scan_into(xv, yv, inclusive=True, init=9)
    np.testing.assert_array_equal(to_host(yv), oracle.inclusive_scan_u32(x, 9))
    check_guards(yb, (off + 1) % 4, n, ys)
    check_guards(xb, off, n, xs)
    w = oracle.generate(n, "u64", 18)
    wb, wv, ws = padded(w, off)
    zb, zv, zs = padded(np.zeros(n, np.uint64), 0)
    mgpu_hip.scan_into(wv, zv, inclusive=False)
    with np.errstate(over="ignore"):
        inc = np.cumsum(w, dtype=np.uint64)
    np.testing.assert_array_equal(to_host(zv), np.concatenate([np.zeros(1, np.uint64), inc[:-1]]))
    check_guards(zb, 0, n, zs)
