"""Graph capture of the plain primitives: radix sort, mergesort, merge and the
scans recorded once into one HIP graph on a side stream, then replayed three
times on fresh inputs (restored by eager copies between replays); every
replay is checked against the oracle."""
import numpy as np
import torch

import mgpu_hip
import oracle
import pytest
from gpu_util import banner, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[mgpu_hip.MERGESORT_TWO_WAY, mgpu_hip.MERGESORT_FOUR_WAY],
                ids=["two_way", "four_way"])
def merge_passes(request):
    old = mgpu_hip.mergesort_passes()
    mgpu_hip.set_mergesort_passes(request.param)
    yield request.param
    mgpu_hip.set_mergesort_passes(old)


def test_primitives_graph_replay(merge_passes):
    banner("primitives graph replay", merge_passes)
    n = (1 << 20) + 3       # several tiles of every kernel, a partial last tile
    na, nb = 300007, 200003
    stream = torch.cuda.Stream()
    src, expect = [], []
    for r in range(3):
        rk = oracle.generate(n, "u32", 60 + r)[np.arange(n) % 5003]   # ties
        mk = oracle.generate(n, "f32_pm1", 70 + r)
        a = np.sort(oracle.generate(na, "u32", 80 + r) % 100000)
        b = np.sort(oracle.generate(nb, "u32", 90 + r) % 100000)
        xi = oracle.generate(n, "i32_0_16", 100 + r)
        xf = oracle.generate(n, "f32_01", 110 + r)
        iota = np.arange(n, dtype=np.uint32)
        src.append([to_dev(v) for v in (rk, iota, mk, iota, a, b, xi, xf)])
        expect.append((oracle.stable_sort_radix(rk, iota, "u32"),
                       oracle.stable_sort(mk, iota, "f32"),
                       oracle.merge(a, b),
                       oracle.exclusive_scan_u32(xi, 7),
                       oracle.inclusive_scan_f64(xf)))
    bufs = [torch.empty_like(t) for t in src[0]]
    rk, rv, mk, mv, a, b, xi, xf = bufs
    c = torch.empty(na + nb, dtype=torch.uint32, device="cuda")
    yi, yf = torch.empty_like(xi), torch.empty_like(xf)

    def restore(r):
        for dst, s in zip(bufs, src[r]):
            dst.copy_(s)

    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        restore(0)
        auxes = (mgpu_hip.radix_sort(rk, rv, stream=stream),
                 mgpu_hip.mergesort_kv(mk, mv, stream=stream),
                 mgpu_hip.merge_keys(a, b, c, "less", stream=stream),
                 mgpu_hip.scan_into(xi, yi, init=7, stream=stream),
                 mgpu_hip.scan_into(xf, yf, inclusive=True, stream=stream))
        stream.synchronize()
        restore(0)
        g = mgpu_hip.Graph(stream)
        with g:
            mgpu_hip.radix_sort(rk, rv, aux=auxes[0], stream=stream)
            mgpu_hip.mergesort_kv(mk, mv, aux=auxes[1], stream=stream)
            mgpu_hip.merge_keys(a, b, c, "less", aux=auxes[2], stream=stream)
            mgpu_hip.scan_into(xi, yi, init=7, aux=auxes[3], stream=stream)
            mgpu_hip.scan_into(xf, yf, inclusive=True, aux=auxes[4], stream=stream)
    stream.synchronize()
    for r in range(3):
        with torch.cuda.stream(stream):
            restore(r)
            g.launch()
        stream.synchronize()
        (ek, ev), (mek, mev), ec, eyi, eyf = expect[r]
        np.testing.assert_array_equal(to_host(rk), ek)
        np.testing.assert_array_equal(to_host(rv), ev)
        np.testing.assert_array_equal(to_host(mk).view(np.uint32), mek.view(np.uint32))
        np.testing.assert_array_equal(to_host(mv), mev)
        np.testing.assert_array_equal(to_host(c), ec)
        np.testing.assert_array_equal(to_host(yi).view(np.uint32), eyi)
        got = to_host(yf).astype(np.float64)
        assert np.max(np.abs(got - eyf) / eyf) <= 2e-6
