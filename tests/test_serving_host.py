"""Host side of the serving engine: request mix, sizes, budget figures and
argument checks (no GPU work)."""
import ctypes

import numpy as np

import mgpu_hip
from mgpu_hip import serving


def test_request_mix_covers_ops_types_and_sizes():
    reqs = serving.make_requests(3000, seed=1)
    ops = {r["op"] for r in reqs}
    kts = {r["key_type"] for r in reqs}
    assert ops == {"sort", "merge", "scan"} and kts == {"u32", "u64", "f32"}
    tot = np.array([serving.elements(r) for r in reqs])
    assert tot.min() >= 1 << 9 and tot.max() <= 1 << 26
    # log-uniform: about half the requests below the geometric midpoint 2^18
    frac = np.mean(tot < (1 << 18))
    assert 0.4 < frac < 0.6
    assert reqs == serving.make_requests(3000, seed=1)


def test_request_bytes_cover_buffers():
    n = 1 << 20
    assert serving.request_bytes("sort", "u32", n) >= 3 * 4 * n
    assert serving.request_bytes("sort", "u64", n) >= 2 * 8 * n + 4 * n
    assert serving.request_bytes("merge", "u32", n) >= 8 * 4 * n
    assert serving.request_bytes("scan", "u64", n) >= 2 * 8 * n
    assert serving.request_bytes("scan", "u32", n) < serving.request_bytes("scan", "u64", n)


def test_server_rejects_bad_arguments():
    lib = mgpu_hip.load()
    serving._bind(lib)
    h = ctypes.c_void_p(0)
    assert lib.mgpu_server_create(0, 1 << 30, 1, ctypes.byref(h)) == 1
    assert lib.mgpu_server_create(4, 0, 1, ctypes.byref(h)) == 1
    assert lib.mgpu_server_create(4, 1 << 30, 1, None) == 1
    r = serving.Request(0, 0, 16, 0)
    rid = ctypes.c_int64(0)
    assert lib.mgpu_server_submit(None, ctypes.byref(r), ctypes.byref(rid)) == 1
    assert lib.mgpu_server_drain(None) == 1
    assert lib.mgpu_server_release(None, 0) == 1
    assert lib.mgpu_server_poll(None, None, 0) == 0


def test_summary_statistics():
    reqs = [{"op": "sort", "key_type": "u32", "n": 1000, "seed": 1},
            {"op": "merge", "key_type": "u32", "n": 500, "seed": 2}]
    resps = [{"index": 0, "arrival_ms": 0.0, "t_done_ms": 1.0},
             {"index": 1, "arrival_ms": 0.5, "t_done_ms": 2.0}]
    s = serving.summarize(resps, reqs)
    assert s["requests"] == 2 and abs(s["makespan_ms"] - 2.0) < 1e-12
    assert abs(s["keys_per_s"] - 2000 / 2e-3) < 1e-6
    assert abs(s["p50_ms"] - 1.25) < 1e-12
