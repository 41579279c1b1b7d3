"""GPU: the serving engine under mixed load.  Every response is checked
against an independent device computation (stable torch.sort for sorts and
merges, integer / float64 prefix sums for scans), and each request's result
is the same run alone as in the mix: bit-identical for sorts, merges and
integer scans, within the f32 scan bound for f32 scans (the carry order of a
look-back scan depends on timing)."""
import numpy as np
import pytest
import torch

from gpu_util import banner
from mgpu_hip import serving

pytestmark = pytest.mark.gpu


def _result_bytes(resp):
    kt, n = resp["key_type"], resp["n"]
    m = 2 * n if resp["op"] == "merge" else n
    keys = serving.device_view(resp["keys"], m, kt).view(torch.uint8).cpu().numpy().copy()
    vals = (serving.device_view(resp["vals"], m, "u32").cpu().numpy().copy()
            if resp["op"] != "scan" else None)
    return keys, vals


def test_serving_mixed_requests_verified():
    banner("serving mixed")
    reqs = serving.make_requests(60, seed=3, lo_log2=10, hi_log2=22)
    sv = serving.Server(slots=8, budget_bytes=16 << 30)
    try:
        resps = serving.run_load(sv, reqs)
        assert len(resps) == len(reqs)
        ops = set()
        for r in resps:
            serving.verify(r)
            ops.add((r["op"], r["key_type"]))
            sv.release(r["id"])
        assert len(ops) >= 7, ops
        st = sv.stats()
        assert st["in_use"] == 0 and st["pending"] == 0 and st["in_flight"] == 0
        assert st["peak"] <= 16 << 30
    finally:
        sv.close()


def test_serving_alone_equals_in_mix():
    banner("serving alone vs mix")
    reqs = serving.make_requests(24, seed=5, lo_log2=10, hi_log2=21)
    # edge sizes: a single element and one tile plus one
    reqs += [{"op": "sort", "key_type": "u64", "n": 1, "seed": 91},
             {"op": "scan", "key_type": "f32", "n": 4097, "seed": 92},
             {"op": "merge", "key_type": "f32", "n": 1, "seed": 93}]
    alone = []
    sv = serving.Server(slots=1, budget_bytes=8 << 30)
    try:
        for rq in reqs:
            rid = sv.submit(rq["op"], rq["key_type"], rq["n"], rq["seed"])
            sv.drain()
            (r,) = sv.poll()
            assert r["id"] == rid
            serving.verify(r)
            alone.append(_result_bytes(r))
            sv.release(rid)
    finally:
        sv.close()
    sv = serving.Server(slots=8, budget_bytes=8 << 30)
    try:
        resps = serving.run_load(sv, reqs)
        for r in resps:
            rq = reqs[r["index"]]
            serving.verify(r)
            keys, vals = _result_bytes(r)
            ak, av = alone[r["index"]]
            if rq["op"] == "scan" and rq["key_type"] == "f32":
                a = ak.view(np.float32).astype(np.float64)
                b = keys.view(np.float32).astype(np.float64)
                nz = a > 0
                assert np.all(np.abs(a - b)[nz] <= 2 * serving.F32_SCAN_RTOL * a[nz])
            else:
                np.testing.assert_array_equal(keys, ak)
                if vals is not None:
                    np.testing.assert_array_equal(vals, av)
            sv.release(r["id"])
    finally:
        sv.close()


def test_serving_budget_admission_and_release():
    banner("serving budget")
    n = 1 << 22
    one = serving.request_bytes("sort", "u32", n)
    budget = 2 * one + one // 2          # two requests fit, a third does not
    sv = serving.Server(slots=8, budget_bytes=budget)
    try:
        ids = [sv.submit("sort", "u32", n, 40 + i) for i in range(5)]
        done = []
        while len(done) < 5:
            rc = sv.lib.mgpu_server_drain(sv.h)
            got = sv.poll()
            assert sv.stats()["in_use"] <= budget
            for r in got:
                serving.verify(r)
                done.append(r["id"])
                sv.release(r["id"])
            if not got:
                assert rc in (0, 2)  # MGPU_OK or MGPU_AUX_TOO_SMALL (budget held)
        assert sorted(done) == sorted(ids)
        assert sv.stats()["peak"] <= budget
        assert sv.stats()["in_use"] == 0
    finally:
        sv.close()


def test_serving_without_kept_inputs():
    banner("serving keep_inputs=0")
    sv = serving.Server(slots=4, budget_bytes=4 << 30, keep_inputs=False)
    try:
        rid = sv.submit("scan", "u32", 100000, 5)
        sv.drain()
        (r,) = sv.poll()
        assert r["id"] == rid and r["status"] == 0
        assert r["input_a"] is None and r["input_b"] is None
        y = serving.device_view(r["keys"], 100000, "u32").view(torch.int32).to(torch.int64)
        assert int(y[0]) == 0
        sv.release(rid)
    finally:
        sv.close()


def test_serving_destroy_with_blocked_pending_requests():
    """Results held unreleased fill the budget, so later requests can never be
    admitted; destroying the server must still return (dropping them)."""
    banner("serving destroy with blocked pending")
    import time
    n = 1 << 20
    one = serving.request_bytes("sort", "u32", n)
    sv = serving.Server(slots=4, budget_bytes=one + one // 2)
    ids = [sv.submit("sort", "u32", n, 70 + i) for i in range(4)]
    rc = sv.lib.mgpu_server_drain(sv.h)
    assert rc == 2  # MGPU_AUX_TOO_SMALL: one result holds the budget, three wait
    got = sv.poll()
    assert len(got) == 1 and got[0]["id"] == ids[0]
    serving.verify(got[0])
    assert sv.stats()["pending"] == 3
    t0 = time.time()
    sv.close()
    assert time.time() - t0 < 10
