"""GPU: calls enqueued on a side stream with stream= and no stream context.

The Python mirror allocates a call's aux on the stream the call runs on, so a
caller that drops the returned aux and goes on allocating on torch's current
stream cannot be handed the aux memory while the library still writes it.
Each case below drops the aux, then immediately allocates and fills tensors
of the aux's size on the default stream before synchronising."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev

pytestmark = pytest.mark.gpu


def _clobber(nbytes, rounds=3):
    """Allocate and fill tensors on the current (default) stream."""
    keep = []
    for r in range(rounds):
        t = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        t.fill_(0xA5 ^ r)
        keep.append(t)
    return keep


def test_radix_sort_side_stream_dropped_aux():
    n = (1 << 24) + 5
    banner("radix_sort stream= without context, aux dropped", n)
    keys = oracle.generate(n, "u32", 11)
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.stable_sort_radix(keys, vals, "u32")
    s = torch.cuda.Stream()
    for it in range(3):
        dk, dv = to_dev(keys), to_dev(vals)
        torch.cuda.synchronize()
        mgpu_hip.radix_sort(dk, dv, stream=s)  # aux returned and dropped
        junk = _clobber(8 * n + (1 << 20))
        s.synchronize()
        torch.cuda.synchronize()
        assert np.array_equal(dk.cpu().numpy(), ek), "keys differ on round %d" % it
        assert np.array_equal(dv.cpu().numpy(), ev), "values differ on round %d" % it
        del junk


def test_scan_and_mergesort_side_stream_dropped_aux():
    n = (1 << 23) + 3
    banner("scan_into + mergesort_kv stream= without context, aux dropped", n)
    x = oracle.generate(n, "i32_0_16", 12)
    keys = oracle.generate(n, "u32", 13)
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.stable_sort(keys, vals, "u32")
    s = torch.cuda.Stream()
    for it in range(2):
        dx = to_dev(x)
        dy = torch.empty_like(dx)
        dk, dv = to_dev(keys), to_dev(vals)
        torch.cuda.synchronize()
        mgpu_hip.scan_into(dx, dy, stream=s)
        mgpu_hip.mergesort_kv(dk, dv, stream=s)
        junk = _clobber(8 * n + (1 << 20))
        s.synchronize()
        torch.cuda.synchronize()
        assert np.array_equal(dy.cpu().numpy().view(np.uint32), oracle.exclusive_scan_u32(x))
        assert np.array_equal(dk.cpu().numpy(), ek) and np.array_equal(dv.cpu().numpy(), ev)
        del junk


def test_caller_aux_used_on_side_stream_is_recorded():
    """A caller-supplied aux allocated on the default stream and used on a
    side stream: the mirror records the side stream on it, so freeing it
    right after the call cannot recycle it under the running sort."""
    n = (1 << 24) + 9
    banner("caller aux on default stream used with stream=", n)
    keys = oracle.generate(n, "u32", 14)
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.stable_sort_radix(keys, vals, "u32")
    size = mgpu_hip.aux_size("mgpu_radix_sort_kv_u32_u32", None, None, n)
    s = torch.cuda.Stream()
    dk, dv = to_dev(keys), to_dev(vals)
    aux = torch.empty(size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    mgpu_hip.radix_sort(dk, dv, aux=aux, stream=s)
    del aux
    junk = _clobber(size)
    s.synchronize()
    torch.cuda.synchronize()
    assert np.array_equal(dk.cpu().numpy(), ek) and np.array_equal(dv.cpu().numpy(), ev)
    del junk

