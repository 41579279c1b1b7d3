"""GPU: the serving engine under mixed load.  Every response is checked
against an independent device computation (stable torch.sort for sorts and
merges, integer / float64 prefix sums for scans), and each request's result
is the same run alone as in the mix: bit-identical for sorts, merges and
integer scans, within the f32 scan bound for f32 scans (the carry order of a
look-back scan depends on timing)."""
import numpy as np
import pytest
import torch

from gpu_util import banner, to_dev, to_host
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


def _caller_cases():
    """Requests on caller buffers: (op, key type, options), sizes mixed."""
    cases = []
    i = 0
    for kt in ("u32", "i32", "u64", "f32"):
        for n in (1, 1000, 70001, 1 << 20):
            cases.append(("sort", kt, n, {"values": "given" if i % 3 == 0 else
                                          ("index" if i % 3 == 1 else None),
                                          "in_place": i % 2 == 0}))
            cases.append(("merge", kt, n, {"n_b": [0, 1, 5000, n // 3][i % 4],
                                           "values": i % 2 == 0}))
            cases.append(("scan", kt, n, {"inclusive": kt == "f32" or i % 2 == 1,
                                          "in_place": i % 3 == 0}))
            i += 1
    return cases


_NP = {"u32": np.uint32, "i32": np.int32, "u64": np.uint64, "f32": np.float32}


def _caller_inputs(op, kt, n, opt, seed):
    import oracle
    kind = {"u32": "u32", "i32": "u32", "u64": "u64", "f32": "f32_pm1"}[kt]
    if op == "scan":
        kind = {"u32": "i32_0_16", "i32": "i32_0_16", "u64": "u64", "f32": "f32_01"}[kt]
    a = oracle.generate(n, kind, seed).view(_NP[kt]) if kt != "f32" else oracle.generate(n, kind, seed)
    a = a[np.arange(n) % 977] if op != "scan" else a  # ties
    b = None
    if op == "merge":
        nb = opt["n_b"] or n
        b = oracle.generate(nb, kind, seed + 1)
        b = (b.view(_NP[kt]) if kt != "f32" else b)[np.arange(nb) % 977]
        a, b = np.sort(a), np.sort(b)
    return np.ascontiguousarray(a), (np.ascontiguousarray(b) if b is not None else None)


def _caller_expected(op, kt, a, b, opt):
    import oracle
    if op == "sort":
        vals = np.arange(len(a), dtype=np.uint32) * 3 + 1 if opt["values"] == "given" else \
            np.arange(len(a), dtype=np.uint32)
        k, v = oracle.stable_sort_radix(a, vals, kt)
        return k, (v if opt["values"] else None)
    if op == "merge":
        av = np.arange(len(a), dtype=np.uint32) * 2
        bv = np.arange(len(b), dtype=np.uint32) * 2 + 1
        if opt["values"]:
            return oracle.merge(a, b, av, bv, kt)
        return oracle.merge(a, b, kt=kt), None
    if kt == "f32":
        exact = oracle.exclusive_scan_f64(a) if not opt["inclusive"] else oracle.inclusive_scan_f64(a)
        return exact, None
    if kt == "u64":
        with np.errstate(over="ignore"):
            inc = np.cumsum(a.view(np.uint64), dtype=np.uint64)
        return (inc if opt["inclusive"] else np.concatenate([[np.uint64(0)], inc[:-1]])), None
    f = oracle.inclusive_scan_u32 if opt["inclusive"] else oracle.exclusive_scan_u32
    return f(a).view(_NP[kt]), None


def _caller_submit(sv, case, seed):
    op, kt, n, opt = case
    a, b = _caller_inputs(op, kt, n, opt, seed)
    da = to_dev(a)
    bufs = {"in_a": da}
    if op == "sort":
        out = da if opt["in_place"] else torch.empty_like(da)
        vo = None
        if opt["values"]:
            vo = torch.empty(n, dtype=torch.int32, device="cuda")
            if opt["values"] == "given":
                bufs["vals_a"] = to_dev((np.arange(n, dtype=np.uint32) * 3 + 1).view(np.int32))
        bufs.update(out=out, vals_out=vo)
    elif op == "merge":
        db = to_dev(b)
        out = torch.empty(len(a) + len(b), dtype=da.dtype, device="cuda")
        bufs.update(in_b=db, out=out)
        if opt["values"]:
            bufs["vals_a"] = to_dev((np.arange(len(a), dtype=np.uint32) * 2).view(np.int32))
            bufs["vals_b"] = to_dev((np.arange(len(b), dtype=np.uint32) * 2 + 1).view(np.int32))
            bufs["vals_out"] = torch.empty(len(a) + len(b), dtype=torch.int32, device="cuda")
    else:
        out = da if opt["in_place"] else torch.empty_like(da)
        bufs.update(out=out, inclusive=opt["inclusive"])
    torch.cuda.synchronize()
    rid = sv.submit_buffers(op, kt, **bufs)
    return rid, (a, b, bufs)


def _caller_check(case, a, b, bufs):
    op, kt, n, opt = case
    ek, ev = _caller_expected(op, kt, a, b, opt)
    out = to_host(bufs["out"])
    if op == "scan" and kt == "f32":
        got = out.astype(np.float64)
        nz = ek > 0
        assert np.all(np.abs(got - ek)[nz] <= serving.F32_SCAN_RTOL * ek[nz]), case
        return out, None
    np.testing.assert_array_equal(out.view(np.uint8), np.asarray(ek).view(np.uint8), str(case))
    vals = None
    if ev is not None:
        vals = to_host(bufs["vals_out"]).view(np.uint32)
        np.testing.assert_array_equal(vals, ev, str(case))
    return out, vals


def test_serving_caller_buffers_alone_and_in_mix():
    """Requests that bring their own device buffers, of every op and key type
    and mixed sizes (values given, index values, key-only; in place and out
    of place; merges of unequal inputs; exclusive and inclusive scans), each
    checked against the oracle run alone and in a concurrent mix, and the
    two results equal (bit for bit except the f32 scans, whose carry order
    follows the timing)."""
    cases = _caller_cases()
    banner("serving caller buffers alone", len(cases))
    alone = []
    sv = serving.Server(slots=1, budget_bytes=4 << 30)
    try:
        for i, case in enumerate(cases):
            rid, (a, b, bufs) = _caller_submit(sv, case, 100 + i)
            sv.drain()
            (r,) = sv.poll()
            assert r["id"] == rid and r["status"] == 0, case
            assert r["keys"] == bufs["out"].data_ptr()
            alone.append(_caller_check(case, a, b, bufs))
            sv.release(rid)
        assert sv.stats()["in_use"] == 0
    finally:
        sv.close()
    banner("serving caller buffers in the mix", len(cases))
    sv = serving.Server(slots=8, budget_bytes=4 << 30)
    try:
        held = {}
        for i, case in enumerate(cases):
            rid, data = _caller_submit(sv, case, 100 + i)
            held[rid] = (i, case, data)
        sv.drain()
        got = sv.poll(len(cases) + 8)
        assert len(got) == len(cases)
        for r in got:
            i, case, (a, b, bufs) = held[r["id"]]
            assert r["status"] == 0, case
            out, vals = _caller_check(case, a, b, bufs)
            if not (case[0] == "scan" and case[1] == "f32"):
                np.testing.assert_array_equal(out.view(np.uint8), alone[i][0].view(np.uint8))
                if vals is not None:
                    np.testing.assert_array_equal(vals, alone[i][1])
            sv.release(r["id"])
        assert sv.stats()["in_use"] == 0 and sv.stats()["peak"] <= 4 << 30
    finally:
        sv.close()
