"""Thread safety per stream: four host threads, each with its own stream and
buffers, call radix sort, mergesort, merge and scan concurrently (the C ABI
calls release the GIL); every result is checked against the oracle."""
import threading

import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host

pytestmark = pytest.mark.gpu


def worker(t, n, rounds, errors):
    try:
        stream = torch.cuda.Stream()
        for r in range(rounds):
            seed = 1000 * t + r
            keys = oracle.generate(n, "u32", seed)[np.arange(n) % 7919]
            vals = np.arange(n, dtype=np.uint32)
            fkeys = oracle.generate(n, "f32_pm1", seed + 500)
            a = np.sort(oracle.generate(n // 2, "u32", seed + 600) % 5000)
            b = np.sort(oracle.generate(n - n // 2, "u32", seed + 700) % 5000)
            x = oracle.generate(n, "i32_0_16", seed + 800)
            with torch.cuda.stream(stream):
                dk, dv = to_dev(keys), to_dev(vals)
                fk, fv = to_dev(fkeys), to_dev(vals)
                da, db = to_dev(a), to_dev(b)
                dc = torch.empty(n, dtype=torch.uint32, device="cuda")
                dx = to_dev(x)
                dy = torch.empty_like(dx)
                mgpu_hip.radix_sort(dk, dv, stream=stream)
                mgpu_hip.mergesort_kv(fk, fv, stream=stream)
                mgpu_hip.merge_keys(da, db, dc, "less", stream=stream)
                mgpu_hip.scan_into(dx, dy, stream=stream)
                stream.synchronize()
            ek, ev = oracle.stable_sort_radix(keys, vals, "u32")
            assert np.array_equal(to_host(dk), ek) and np.array_equal(to_host(dv), ev), "radix"
            mk, mv = oracle.stable_sort(fkeys, vals, "f32")
            assert np.array_equal(to_host(fk).view(np.uint32), mk.view(np.uint32)), "mergesort"
            assert np.array_equal(to_host(fv), mv), "mergesort vals"
            assert np.array_equal(to_host(dc), oracle.merge(a, b)), "merge"
            assert np.array_equal(to_host(dy).view(np.uint32), oracle.exclusive_scan_u32(x)), "scan"
    except Exception as e:  # reported by the main thread
        errors.append((t, repr(e)))


def test_concurrent_host_threads():
    banner("host threads")
    errors = []
    threads = [threading.Thread(target=worker, args=(t, 300007 + 1000 * t, 4, errors))
               for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
