"""GPU parity: scan and reduce against the oracle (bit-exact for integers;
float32 inclusive scan within a stated relative error of the exact float64
prefix)."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, dev_equal, gen_dev, to_dev, to_host

pytestmark = pytest.mark.gpu

NV = 256 * 80  # elements per tile of the library's 32-bit integer scan
F32_NV = 256 * 96  # elements per tile of its float32 scan
SIZES = [1, 2, 3, 4, 5, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097, NV - 1, NV, NV + 1,
         3 * NV + 7, 64 * NV + 1, 1 << 20,
         (1 << 22) + 13]

# float32 scan contract: float32 sums only inside aligned 256-element blocks
# (at most 3 thread-serial additions + 6 wave-scan levels + the wave prefix
# = 10 on any path); every sum above a block and the carries between tiles
# are float64, and one rounding at the output.  For non-negative inputs the
# relative error is below gamma_10 + u ~ 6.6e-7; the bound below is 2e-6.
F32_SCAN_RTOL = 2e-6


@pytest.mark.parametrize("n", SIZES)
def test_exclusive_scan_i32_in_place(n):
    banner("scan_exclusive_i32", n)
    x = oracle.generate(n, "i32_0_16", 1)
    d = to_dev(x)
    mgpu_hip.scan(d)
    np.testing.assert_array_equal(to_host(d).view(np.uint32), oracle.exclusive_scan_u32(x))


@pytest.mark.parametrize("n", [1, 4097, 1 << 20])
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_u32_out_of_place_with_init(n, inclusive):
    banner("scan_u32", n, inclusive)
    x = oracle.generate(n, "u32", 2)
    d = to_dev(x)
    out = torch.empty_like(d)
    mgpu_hip.scan_into(d, out, inclusive=inclusive, init=123456789)
    exp = (oracle.inclusive_scan_u32 if inclusive else oracle.exclusive_scan_u32)(x, 123456789)
    np.testing.assert_array_equal(to_host(out), exp)
    np.testing.assert_array_equal(to_host(d), x)  # input untouched


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_scan_misaligned_pointers(offset):
    banner("scan misaligned", offset)
    n = 3 * NV + 5
    x = oracle.generate(n + offset, "i32_0_16", 3)
    d = to_dev(x)
    view = d[offset:]
    mgpu_hip.scan(view)
    got = to_host(d)
    np.testing.assert_array_equal(got[:offset], x[:offset])
    np.testing.assert_array_equal(got[offset:].view(np.uint32), oracle.exclusive_scan_u32(x[offset:]))


def test_scan_int32_wraps_like_uint32():
    banner("scan wrap")
    x = np.full(50000, 2 ** 30 - 7, dtype=np.int32)
    d = to_dev(x)
    mgpu_hip.scan(d)
    np.testing.assert_array_equal(to_host(d).view(np.uint32), oracle.exclusive_scan_u32(x))


def test_vk_scan_known_answer():
    # examples/vk_scan/vk_scan.cxx:17-23: host[i] = i, count 10000.
    banner("vk_scan kat")
    i = np.arange(10000, dtype=np.int32)
    d = to_dev(i)
    mgpu_hip.scan(d)
    exp = (i.astype(np.int64) * (i.astype(np.int64) - 1) // 2)
    exp[0] = 0
    np.testing.assert_array_equal(to_host(d).astype(np.int64), exp)


@pytest.mark.parametrize("n", [1, 4095, 4097, F32_NV - 1, F32_NV, F32_NV + 1, 3 * F32_NV + 7,
                               1 << 20, (1 << 22) + 13])
@pytest.mark.parametrize("inclusive", [True, False])
def test_scan_f32_error_bound(n, inclusive):
    banner("scan_f32", n, inclusive)
    x = oracle.generate(n, "f32_01", 4)
    d = to_dev(x)
    out = torch.empty_like(d)
    mgpu_hip.scan_into(d, out, inclusive=inclusive)
    exact = (oracle.inclusive_scan_f64 if inclusive else oracle.exclusive_scan_f64)(x)
    got = to_host(out).astype(np.float64)
    nz = exact > 0
    rel = np.abs(got[nz] - exact[nz]) / exact[nz]
    assert rel.max(initial=0) <= F32_SCAN_RTOL, rel.max()
    assert np.all(got[~nz] == 0)


def test_scan_determinism_integers():
    banner("scan determinism")
    n = (1 << 22) + 3
    d = gen_dev(n, "i32_0_16", 5)
    a = torch.empty_like(d)
    b = torch.empty_like(d)
    mgpu_hip.scan_into(d, a)
    mgpu_hip.scan_into(d, b)
    assert dev_equal(a, b)


# Reduce tiles hold 8192 four-byte elements (256 threads x 8 x 128-bit).
@pytest.mark.parametrize("n", [1, 3, 4, 1000, 4097, 8191, 8192, 8195, 3 * 8192 + 5, 1 << 20,
                               (1 << 22) + 13])
@pytest.mark.parametrize("offset", [0, 1])
def test_reduce(n, offset):
    banner("reduce", n, offset)
    xi = oracle.generate(n + offset, "u32", 6)
    di = to_dev(xi)
    out = torch.zeros(1, dtype=torch.uint32, device="cuda")
    mgpu_hip.reduce(di[offset:], out, init=5)
    assert int(to_host(out)[0]) == oracle.reduce_u32(xi[offset:], 5)
    xf = oracle.generate(n + offset, "f32_01", 7)
    df = to_dev(xf)
    outf = torch.zeros(1, dtype=torch.float32, device="cuda")
    mgpu_hip.reduce(df[offset:], outf)
    exact = oracle.reduce_f64(xf[offset:])
    # float64 accumulation, one rounding to float32.
    assert abs(float(to_host(outf)[0]) - exact) <= abs(exact) * 2 ** -24 * 1.01


def test_device_generator_matches_oracle():
    banner("generate")
    for kind in ["u32", "lowentropy", "i32_0_16", "f32_01", "u64", "f32_pm1", "iota"]:
        t = gen_dev(100003, kind, 9, 77)
        np.testing.assert_array_equal(to_host(t), oracle.generate(100003, kind, 9, 77))


@pytest.mark.parametrize("limit", [0, 1])
@pytest.mark.parametrize("inplace", [True, False])
def test_scan_forced_lookback_fallback(limit, inplace):
    """A tiny spin limit makes tiles sum late predecessors from their input
    (in place: with the re-check of the published word); results unchanged."""
    banner("scan fallback", limit, inplace)
    n = 300 * NV + 17
    x = oracle.generate(n, "i32_0_16", 21)
    f = oracle.generate(n, "f32_01", 22)
    old = mgpu_hip.lookback_spin_limit()
    mgpu_hip.lookback_fallbacks(reset=True)
    try:
        mgpu_hip.set_lookback_spin_limit(limit)
        d = to_dev(x)
        df = to_dev(f)
        if inplace:
            mgpu_hip.scan(d)
            mgpu_hip.scan_into(df, df, inclusive=True)
            out, outf = d, df
        else:
            out, outf = torch.empty_like(d), torch.empty_like(df)
            mgpu_hip.scan_into(d, out)
            mgpu_hip.scan_into(df, outf, inclusive=True)
        torch.cuda.synchronize()
    finally:
        mgpu_hip.set_lookback_spin_limit(old)
    print("fallbacks", mgpu_hip.lookback_fallbacks(), flush=True)
    np.testing.assert_array_equal(to_host(out).view(np.uint32), oracle.exclusive_scan_u32(x))
    exact = oracle.inclusive_scan_f64(f)
    assert np.max(np.abs(to_host(outf).astype(np.float64) - exact) / exact) <= F32_SCAN_RTOL


U64_NV = 256 * 48  # tile of the 64-bit scan above 2^20 elements
U64_SMALL_NV = 256 * 8  # its tile up to 2^20 elements
U64_SMALL_MAX = 1 << 20


def u64_expected(x, inclusive, init):
    with np.errstate(over="ignore"):
        inc = np.cumsum(x, dtype=np.uint64) + np.uint64(init)
    return inc if inclusive else np.concatenate([np.array([init], np.uint64), inc[:-1]])


@pytest.mark.parametrize("n", [1, 7, U64_SMALL_NV - 1, U64_SMALL_NV, U64_SMALL_NV + 1,
                               U64_NV - 1, U64_NV, U64_NV + 1, 3 * U64_NV + 5,
                               U64_SMALL_MAX, U64_SMALL_MAX + 1, 300 * U64_NV + 1, (1 << 22) + 1])
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_u64(n, inclusive):
    banner("scan u64", n, inclusive)
    x = oracle.generate(n, "u64", 30)
    d = to_dev(x)
    out = torch.empty_like(d)
    mgpu_hip.scan_into(d, out, inclusive=inclusive, init=12345)
    np.testing.assert_array_equal(to_host(out), u64_expected(x, inclusive, 12345))
    r = torch.zeros(1, dtype=torch.uint64, device="cuda")
    mgpu_hip.reduce(d, r, init=5)
    with np.errstate(over="ignore"):
        assert int(to_host(r)[0]) == int(np.sum(x, dtype=np.uint64) + np.uint64(5))


@pytest.mark.parametrize("limit", [0, 1, 512])
@pytest.mark.parametrize("in_place", [False, True])
@pytest.mark.parametrize("n", [300 * U64_SMALL_NV + 5, 700 * U64_NV + 17])
def test_scan_u64_lookback_fallback_and_in_place(limit, in_place, n):
    """A tiny spin limit makes tiles sum late predecessors from their input
    (in place: with the re-check of the predecessor's words); the result is
    unchanged, with the small tiles and the large ones."""
    banner("scan u64 fallback", limit, in_place, n)
    x = oracle.generate(n, "u64", 31)
    old = mgpu_hip.lookback_spin_limit()
    try:
        mgpu_hip.set_lookback_spin_limit(limit)
        d = to_dev(x)
        out = d if in_place else torch.empty_like(d)
        mgpu_hip.scan_into(d, out, inclusive=False, init=3)
        torch.cuda.synchronize()
    finally:
        mgpu_hip.set_lookback_spin_limit(old)
    np.testing.assert_array_equal(to_host(out), u64_expected(x, False, 3))