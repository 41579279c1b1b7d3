"""GPU: mgpu_scan_i32_noalloc (status words in a fixed device array, no
allocation) against mgpu_scan_i32 (caller aux) through the C ABI, bit-exact,
on consecutive calls of decreasing size; its size limit is rejected before
any launch."""
import ctypes

import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, gen_dev

pytestmark = pytest.mark.gpu

MAX_COUNT = 131072 * 256 * 80  # status words of 131072 tiles of 20480 elements


def _noalloc():
    lib = mgpu_hip.load()
    f = lib.mgpu_scan_i32_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                  ctypes.c_int, ctypes.c_int32]
    return f


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("inclusive", [False, True])
def test_noalloc_scan_matches_aux_scan(inclusive):
    f = _noalloc()
    for n, seed in (((1 << 27) + 12345, 1), (100003, 2), (20480 * 3 + 1, 3)):
        banner("scan_i32_noalloc", n, "inclusive" if inclusive else "exclusive", seed)
        x = gen_dev(n, "i32_0_16", seed)
        ref = torch.empty_like(x)
        mgpu_hip.scan_into(x, ref, inclusive=inclusive, init=7)
        out = torch.full_like(x, -1)
        st = f(_stream(), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), n,
               mgpu_hip.INCLUSIVE if inclusive else mgpu_hip.EXCLUSIVE, 7)
        assert st == mgpu_hip.MGPU_OK
        torch.cuda.synchronize()
        assert torch.equal(out, ref), "noalloc scan differs from the aux scan at n=%d" % n
        if n < (1 << 20):
            exp = (oracle.inclusive_scan_u32(x.cpu().numpy(), 7) if inclusive
                   else oracle.exclusive_scan_u32(x.cpu().numpy(), 7))
            assert np.array_equal(out.cpu().numpy().view(np.uint32), exp)


def test_noalloc_scan_in_place_and_limits():
    f = _noalloc()
    n = 1 << 22
    banner("scan_i32_noalloc in place", n)
    x = gen_dev(n, "i32_0_16", 4)
    ref = torch.empty_like(x)
    mgpu_hip.scan_into(x, ref)
    st = f(_stream(), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()), n,
           mgpu_hip.EXCLUSIVE, 0)
    assert st == mgpu_hip.MGPU_OK
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    # Rejected before any launch: beyond the scratch, bad type, negative count.
    small = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = ctypes.c_void_p(small.data_ptr())
    assert f(_stream(), p, p, MAX_COUNT + 1, mgpu_hip.EXCLUSIVE, 0) == mgpu_hip.MGPU_UNSUPPORTED
    assert f(_stream(), p, p, 16, 5, 0) == mgpu_hip.MGPU_INVALID_ARG
    assert f(_stream(), p, p, -1, mgpu_hip.EXCLUSIVE, 0) == mgpu_hip.MGPU_INVALID_ARG
    assert f(_stream(), p, p, 0, mgpu_hip.EXCLUSIVE, 0) == mgpu_hip.MGPU_OK
    torch.cuda.synchronize()
    assert bool((small == 0).all())
