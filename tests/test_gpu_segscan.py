"""GPU parity: segmented scan against the oracle's per-segment std::scan
(bit-exact for i32/u32 modulo 2^32; f32 within the scan's relative-error
bound of the exact float64 per-segment prefix), over segment layouts that
put starts on, next to and far from tile boundaries, empty segments, one
segment per element, and segments spanning many tiles (long look-back)."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host
from mgpu_hip import workloads

pytestmark = pytest.mark.gpu

NV_INT = 256 * 48  # elements per tile of the library's segmented scan: integers
NV_F32 = 256 * 32  # float32
F32_SCAN_RTOL = 2e-6
SIZES = [1, 5, 4095, 4096, 4097, 3 * NV_INT + 7, 3 * NV_F32 + 7, 1 << 20, (1 << 22) + 13]
LAYOUTS = ["none", "single0", "every", "lognormal", "tile_edges", "sparse", "dupes"]


def layout(name, n, seed=0, nv=NV_INT):
    if name == "none":
        return np.zeros(0, np.int32)
    if name == "single0":
        return np.zeros(1, np.int32)
    if name == "every":
        return np.arange(n, dtype=np.int32)
    if name == "lognormal":
        lens = workloads.lognormal_lengths(max(1, n // 200), 256.0, 1.5, 1, 1 << 20, seed=seed)
        s = workloads.segment_starts(lens)
        return s[s < n]
    if name == "tile_edges":
        t = np.arange(0, n + nv, nv, dtype=np.int64)
        s = np.unique(np.concatenate([t - 1, t, t + 1]))
        return s[(s >= 0) & (s < n)].astype(np.int32)
    if name == "sparse":  # few starts, segments over many tiles
        return np.array([0, n // 3, n // 3 + 1, (2 * n) // 3], dtype=np.int32)
    if name == "dupes":  # empty segments, starts past the end
        s = np.array([0, 7, 7, 7, nv, nv, n // 2, n, n, n + 9], dtype=np.int64)
        return np.sort(np.clip(s, 0, n + 9)).astype(np.int32)
    raise ValueError(name)


def check(kind, n, segs, inclusive, init, in_place=False):
    x = oracle.generate(n, kind, 5)
    d_in = to_dev(x)
    d_out = d_in if in_place else torch.empty_like(d_in)
    mgpu_hip.segmented_scan(d_in, d_out, to_dev(segs), inclusive=inclusive, init=init)
    got = to_host(d_out)
    if kind == "f32_01":
        exact = oracle.segmented_scan_f64(x, segs, inclusive, float(init))
        g = got.astype(np.float64)
        zero = exact == 0
        assert np.all(g[zero] == 0)
        rel = np.abs(g[~zero] - exact[~zero]) / exact[~zero]
        assert rel.size == 0 or rel.max() <= F32_SCAN_RTOL, rel.max()
    else:
        exp = oracle.segmented_scan_u32(x, segs, inclusive, init)
        np.testing.assert_array_equal(got.view(np.uint32), exp)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("lay", LAYOUTS)
def test_segscan_i32_exclusive(n, lay):
    banner("segscan i32 exc", n, lay)
    check("i32_0_16", n, layout(lay, n, 1), False, 3)


@pytest.mark.parametrize("n", [4097, 1 << 20])
@pytest.mark.parametrize("lay", LAYOUTS)
def test_segscan_u32_inclusive_wraps(n, lay):
    banner("segscan u32 inc", n, lay)
    check("u32", n, layout(lay, n, 2), True, 0xFFFFFFF0)


@pytest.mark.parametrize("n", [4095, 3 * NV_F32 + 7, (1 << 22) + 13])
@pytest.mark.parametrize("lay", LAYOUTS)
@pytest.mark.parametrize("inclusive", [False, True])
def test_segscan_f32_error_bound(n, lay, inclusive):
    banner("segscan f32", n, lay, inclusive)
    check("f32_01", n, layout(lay, n, 3, NV_F32), inclusive, 0.0)


@pytest.mark.parametrize("lay", ["lognormal", "sparse", "tile_edges"])
def test_segscan_in_place(lay):
    banner("segscan in place", lay)
    n = (1 << 21) + 77
    check("i32_0_16", n, layout(lay, n, 4), False, 0, in_place=True)
    check("f32_01", n, layout(lay, n, 4, NV_F32), True, 0.0, in_place=True)


@pytest.mark.parametrize("in_place", [False, True])
def test_segscan_forced_lookback_fallback(in_place):
    """Spin limit 0: every missing predecessor is summed from its input
    (in place: with the re-check of its word); results are unchanged."""
    banner("segscan fallback", in_place)
    old = mgpu_hip.lookback_spin_limit()
    mgpu_hip.set_lookback_spin_limit(0)
    try:
        n = (1 << 22) + 13
        for lay in ("sparse", "lognormal", "none"):
            check("i32_0_16", n, layout(lay, n, 6), False, 1, in_place=in_place)
            check("f32_01", n, layout(lay, n, 6, NV_F32), True, 0.0, in_place=in_place)
    finally:
        mgpu_hip.set_lookback_spin_limit(old)


def test_segscan_graph_replay():
    banner("segscan graph replay")
    n = (1 << 20) + 3
    segs = to_dev(layout("lognormal", n, 7))
    stream = torch.cuda.Stream()
    srcs = [to_dev(oracle.generate(n, "i32_0_16", 30 + r)) for r in range(3)]
    d_in = torch.empty(n, dtype=torch.int32, device="cuda")
    d_out = torch.empty_like(d_in)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        d_in.copy_(srcs[0])
        aux = mgpu_hip.segmented_scan(d_in, d_out, segs, stream=stream)
        stream.synchronize()
        g = mgpu_hip.Graph(stream)
        with g:
            mgpu_hip.segmented_scan(d_in, d_out, segs, aux=aux, stream=stream)
    stream.synchronize()
    hs = to_host(segs)
    for r in range(3):
        with torch.cuda.stream(stream):
            d_in.copy_(srcs[r])
            g.launch()
        stream.synchronize()
        exp = oracle.segmented_scan_u32(to_host(srcs[r]), hs, False, 0)
        np.testing.assert_array_equal(to_host(d_out).view(np.uint32), exp)


def test_segscan_full_size_i32_65536_segments():
    """2^29 int32 in [0, 16) over 65536 lognormal segments, checked on the
    device with int64 prefix sums (exact) reduced modulo 2^32."""
    banner("segscan full size")
    n = 1 << 29
    lens = workloads.lognormal_lengths(65536, 2048.0, 1.5, 1, 1 << 22, seed=0)
    starts = workloads.segment_starts(lens)
    starts = starts[starts < n]
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    mgpu_hip.generate(x, "i32_0_16", 0)
    out = torch.empty_like(x)
    segs = to_dev(starts)
    mgpu_hip.segmented_scan(x, out, segs)
    heads = torch.zeros(n, dtype=torch.bool, device="cuda")
    heads[segs.long()] = True
    heads[0] = True
    c = torch.cumsum(x.long(), 0) - x.long()           # exclusive global prefix
    seg_id = torch.cumsum(heads.int(), 0) - 1
    base = c[torch.nonzero(heads).squeeze(1)][seg_id]
    exp = (c - base) & 0xFFFFFFFF
    del c, seg_id, base, heads
    assert torch.equal(out.long() & 0xFFFFFFFF, exp)
