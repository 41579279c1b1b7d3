"""Client side of the serving engine (mgpu_server_* in the C ABI).

Server wraps the engine; run_load drives it with Poisson arrivals of mixed
sort / merge / scan requests (sizes log-uniform, key types u32 / u64 / f32)
and verify() checks every response on the device against independent torch
computations: stable torch.sort for sorts and merges (bit-exact), int64 /
float64 prefix sums for scans (u32 bit-exact modulo 2^32; f32 within the
scan's relative-error bound), numpy uint64 prefix sums for u64 scans.
"""
import ctypes
import time

import numpy as np
import torch

from . import _check, load, MgpuError
from . import workloads

SORT, MERGE, SCAN = 0, 1, 2
OPS = {"sort": SORT, "merge": MERGE, "scan": SCAN}
KEY = {"u32": 0, "i32": 1, "u64": 2, "f32": 3}
KEY_NAME = {v: k for k, v in KEY.items()}
OP_NAME = {v: k for k, v in OPS.items()}
F32_SCAN_RTOL = 2e-6


class Request(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("key_type", ctypes.c_int32), ("n", ctypes.c_int64),
                ("seed", ctypes.c_uint64)]


class RequestBuffers(ctypes.Structure):
    _fields_ = [("in_a", ctypes.c_void_p), ("in_b", ctypes.c_void_p),
                ("vals_a", ctypes.c_void_p), ("vals_b", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("vals_out", ctypes.c_void_p),
                ("n_b", ctypes.c_int64), ("scan_type", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Response(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int32), ("op", ctypes.c_int32),
                ("key_type", ctypes.c_int32), ("slot", ctypes.c_int32), ("n", ctypes.c_int64),
                ("keys", ctypes.c_void_p), ("vals", ctypes.c_void_p),
                ("input_a", ctypes.c_void_p), ("input_b", ctypes.c_void_p),
                ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double)]


def _bind(lib):
    if getattr(lib, "_serving_bound", False):
        return
    vp = ctypes.c_void_p
    lib.mgpu_server_create.restype = ctypes.c_int
    lib.mgpu_server_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(vp)]
    lib.mgpu_server_now.restype = ctypes.c_double
    lib.mgpu_server_now.argtypes = [vp]
    lib.mgpu_server_request_bytes.restype = ctypes.c_size_t
    lib.mgpu_server_request_bytes.argtypes = [ctypes.POINTER(Request)]
    lib.mgpu_server_submit.restype = ctypes.c_int
    lib.mgpu_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]
    lib.mgpu_server_submit_buffers.restype = ctypes.c_int
    lib.mgpu_server_submit_buffers.argtypes = [vp, ctypes.POINTER(Request),
                                               ctypes.POINTER(RequestBuffers),
                                               ctypes.POINTER(ctypes.c_int64)]
    lib.mgpu_server_poll.restype = ctypes.c_int
    lib.mgpu_server_poll.argtypes = [vp, ctypes.POINTER(Response), ctypes.c_int]
    lib.mgpu_server_drain.restype = ctypes.c_int
    lib.mgpu_server_drain.argtypes = [vp]
    lib.mgpu_server_release.restype = ctypes.c_int
    lib.mgpu_server_release.argtypes = [vp, ctypes.c_int64]
    lib.mgpu_server_stats.restype = None
    lib.mgpu_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]
    lib.mgpu_server_destroy.restype = None
    lib.mgpu_server_destroy.argtypes = [vp]
    lib._serving_bound = True


def request_bytes(op, key_type, n):
    lib = load()
    _bind(lib)
    r = Request(OPS[op] if isinstance(op, str) else op,
                KEY[key_type] if isinstance(key_type, str) else key_type, n, 0)
    return int(lib.mgpu_server_request_bytes(ctypes.byref(r)))


class Server:
    def __init__(self, slots=8, budget_bytes=64 << 30, keep_inputs=True):
        self.lib = load()
        _bind(self.lib)
        self.h = ctypes.c_void_p(0)
        _check(self.lib.mgpu_server_create(slots, budget_bytes, 1 if keep_inputs else 0,
                                           ctypes.byref(self.h)), "server_create")
        self.budget = budget_bytes

    def now(self):
        return self.lib.mgpu_server_now(self.h)

    def submit(self, op, key_type, n, seed):
        r = Request(OPS[op] if isinstance(op, str) else op,
                    KEY[key_type] if isinstance(key_type, str) else key_type, n, seed)
        rid = ctypes.c_int64(-1)
        _check(self.lib.mgpu_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)),
               "server_submit")
        return rid.value

    def submit_buffers(self, op, key_type, in_a, out, in_b=None, vals_a=None, vals_b=None,
                       vals_out=None, inclusive=False):
        """A request on caller-owned device tensors (kept alive and unmodified
        by the caller until the response is polled); see mgpu_server_submit_buffers."""
        for t in (in_a, out, in_b, vals_a, vals_b, vals_out):
            if t is not None and (not t.is_cuda or not t.is_contiguous()):
                raise ValueError("buffers must be contiguous device tensors")
        n = in_a.numel()
        nb = in_b.numel() if in_b is not None else 0
        m = n + nb if op == "merge" else n
        if out.numel() < m or out.element_size() != in_a.element_size():
            raise ValueError("out must hold %d keys of the input's size" % m)
        if vals_out is not None and (vals_out.numel() < m or vals_out.element_size() != 4):
            raise ValueError("vals_out must hold %d 32-bit values" % m)
        for v, cnt in ((vals_a, n), (vals_b, nb)):
            if v is not None and (v.numel() != cnt or v.element_size() != 4):
                raise ValueError("values must be 32-bit, one per key")
        if in_b is not None and in_b.element_size() != in_a.element_size():
            raise ValueError("a and b keys must have one size")
        ptr = lambda t: t.data_ptr() if t is not None else None
        r = Request(OPS[op], KEY[key_type], n, 0)
        b = RequestBuffers(ptr(in_a), ptr(in_b), ptr(vals_a), ptr(vals_b), ptr(out), ptr(vals_out),
                           nb, 1 if inclusive else 0, 0)
        rid = ctypes.c_int64(-1)
        _check(self.lib.mgpu_server_submit_buffers(self.h, ctypes.byref(r), ctypes.byref(b),
                                                   ctypes.byref(rid)), "server_submit_buffers")
        return rid.value

    def poll(self, max_n=256):
        buf = (Response * max_n)()
        n = self.lib.mgpu_server_poll(self.h, buf, max_n)
        return [_resp_dict(buf[i]) for i in range(n)]

    def drain(self):
        _check(self.lib.mgpu_server_drain(self.h), "server_drain")

    def release(self, rid):
        _check(self.lib.mgpu_server_release(self.h, rid), "server_release")

    def stats(self):
        in_use, peak = ctypes.c_size_t(0), ctypes.c_size_t(0)
        pend, fl = ctypes.c_int64(0), ctypes.c_int64(0)
        self.lib.mgpu_server_stats(self.h, ctypes.byref(in_use), ctypes.byref(peak),
                                   ctypes.byref(pend), ctypes.byref(fl))
        return {"in_use": in_use.value, "peak": peak.value, "pending": pend.value,
                "in_flight": fl.value}

    def close(self):
        if self.h:
            self.lib.mgpu_server_destroy(self.h)
            self.h = ctypes.c_void_p(0)

    def __del__(self):
        self.close()


def _resp_dict(r):
    return {"id": r.id, "status": r.status, "op": OP_NAME[r.op], "key_type": KEY_NAME[r.key_type],
            "slot": r.slot, "n": r.n, "keys": r.keys, "vals": r.vals, "input_a": r.input_a,
            "input_b": r.input_b, "t_submit_ms": r.t_submit_ms, "t_start_ms": r.t_start_ms,
            "t_done_ms": r.t_done_ms}


# ---------------------------------------------------------------------------
# Device views of response buffers (the server owns them until release)

_TORCH = {"u32": torch.uint32, "i32": torch.int32, "u64": torch.uint64, "f32": torch.float32}
_BYTES = {"u32": 4, "i32": 4, "u64": 8, "f32": 4}


def device_view(ptr, n, kt):
    """A torch tensor over n elements of device memory owned by the server
    (no ownership taken, valid until the request is released); where torch
    cannot wrap a foreign pointer, a device-to-device copy instead."""
    if n == 0:
        return torch.empty(0, dtype=_TORCH[kt], device="cuda")
    nbytes = n * _BYTES[kt]
    try:
        u8 = torch.as_tensor(_CudaArray(ptr, nbytes), device="cuda")
    except (TypeError, RuntimeError, ValueError):
        u8 = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        if _hip().hipMemcpy(ctypes.c_void_p(u8.data_ptr()), ctypes.c_void_p(ptr),
                            ctypes.c_size_t(nbytes), 3) != 0:
            raise RuntimeError("hipMemcpy of a response buffer failed")
    return u8.view(_TORCH[kt])


_HIP = []


def _hip():
    if not _HIP:
        _HIP.append(ctypes.CDLL("libamdhip64.so"))
    return _HIP[0]


class _CudaArray:
    def __init__(self, ptr, nbytes):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1",
                                         "data": (ptr, False), "version": 3}


def _signed_order(t):
    if t.dtype == torch.uint32:
        return t.view(torch.int32) ^ torch.tensor(-2 ** 31, dtype=torch.int32, device=t.device)
    if t.dtype == torch.uint64:
        return t.view(torch.int64) ^ torch.tensor(-2 ** 63, dtype=torch.int64, device=t.device)
    return t


def verify(resp):
    """Check one response against an independent computation; returns the
    maximum relative error for f32 scans (0.0 otherwise) or raises."""
    kt, n, op = resp["key_type"], resp["n"], resp["op"]
    if resp["status"] != 0:
        raise MgpuError(resp["status"], "request %d" % resp["id"])
    if op == "sort":
        keys_in = device_view(resp["input_a"], n, kt)
        ks = device_view(resp["keys"], n, kt)
        vs = device_view(resp["vals"], n, "u32")
        order = torch.sort(_signed_order(keys_in), stable=True)
        if not torch.equal(order.values, _signed_order(ks)):
            raise AssertionError("sort %d: keys differ" % resp["id"])
        if not torch.equal(order.indices.to(torch.int32), vs.view(torch.int32)):
            raise AssertionError("sort %d: values differ" % resp["id"])
        return 0.0
    if op == "merge":
        a = device_view(resp["input_a"], n, kt)
        b = device_view(resp["input_b"], n, kt)
        ks = device_view(resp["keys"], 2 * n, kt)
        vs = device_view(resp["vals"], 2 * n, "u32")
        cat = torch.cat([_signed_order(a), _signed_order(b)])
        order = torch.sort(cat, stable=True)
        if not torch.equal(order.values, _signed_order(ks)):
            raise AssertionError("merge %d: keys differ" % resp["id"])
        if not torch.equal(order.indices.to(torch.int32), vs.view(torch.int32)):
            raise AssertionError("merge %d: values differ" % resp["id"])
        return 0.0
    x = device_view(resp["input_a"], n, kt)
    y = device_view(resp["keys"], n, kt)
    if kt == "u64":
        xh = x.view(torch.int64).cpu().numpy().view(np.uint64)
        with np.errstate(over="ignore"):
            inc = np.cumsum(xh, dtype=np.uint64)
        exc = np.concatenate([np.zeros(1, np.uint64), inc[:-1]])
        if not np.array_equal(y.view(torch.int64).cpu().numpy().view(np.uint64), exc):
            raise AssertionError("scan %d: u64 differs" % resp["id"])
        return 0.0
    if kt in ("u32", "i32"):
        inc = torch.cumsum(x.view(torch.int32).to(torch.int64) & 0xFFFFFFFF, 0)
        exc = torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), inc[:-1]]) & 0xFFFFFFFF
        if not torch.equal(exc, y.view(torch.int32).to(torch.int64) & 0xFFFFFFFF):
            raise AssertionError("scan %d: u32 differs" % resp["id"])
        return 0.0
    exact = torch.cumsum(x.double(), 0)  # exact: multiples of 2^-24, prefix < 2^53
    nz = exact > 0
    rel = ((y.double() - exact).abs()[nz] / exact[nz]).max().item() if bool(nz.any()) else 0.0
    if rel > F32_SCAN_RTOL:
        raise AssertionError("scan %d: f32 rel err %g" % (resp["id"], rel))
    return rel


# ---------------------------------------------------------------------------
# Load generation

def make_requests(count, seed=0, lo_log2=10, hi_log2=26, key_types=("u32", "u64", "f32"),
                  ops=("sort", "merge", "scan")):
    """count requests: op and key type uniform, total elements log-uniform
    over [2^lo, 2^hi] (a merge takes two inputs of half the size)."""
    u = workloads.splitmix_u64(seed, 5, count)
    sizes = workloads.loguniform_sizes(count, lo_log2, hi_log2, seed)
    reqs = []
    for i in range(count):
        op = ops[int(u[i] % np.uint64(len(ops)))]
        kt = key_types[int((u[i] >> np.uint64(8)) % np.uint64(len(key_types)))]
        n = int(sizes[i])
        if op == "merge":
            n = max(1, n // 2)
        reqs.append({"op": op, "key_type": kt, "n": n, "seed": seed * 100003 + i + 1})
    return reqs


def elements(req):
    return req["n"] * (2 if req["op"] == "merge" else 1)


def run_load(server, reqs, rate=None, poll_every=16):
    """Submit reqs at Poisson arrival times (rate per second; None = all at
    once), then drain.  Returns the responses (with 'arrival_ms')."""
    arrivals = (workloads.poisson_arrivals(len(reqs), rate, seed=7) * 1e3
                if rate else np.zeros(len(reqs)))
    t0 = server.now() + 2.0
    ids = {}
    out = []
    for i, rq in enumerate(reqs):
        target = t0 + arrivals[i]
        while server.now() < target:
            pass
        rid = server.submit(rq["op"], rq["key_type"], rq["n"], rq["seed"])
        ids[rid] = (i, target)
        if i % poll_every == 0:
            out.extend(server.poll())
    server.drain()
    while True:
        got = server.poll()
        if not got:
            break
        out.extend(got)
    for r in out:
        i, target = ids[r["id"]]
        r["index"] = i
        r["arrival_ms"] = target
    out.sort(key=lambda r: r["index"])
    return out


def summarize(resps, reqs):
    lat = np.array([r["t_done_ms"] - r["arrival_ms"] for r in resps])
    first = min(r["arrival_ms"] for r in resps)
    last = max(r["t_done_ms"] for r in resps)
    elems = sum(elements(reqs[r["index"]]) for r in resps)
    return {"requests": len(resps), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "makespan_ms": last - first, "keys_per_s": elems / ((last - first) / 1e3),
            "requests_per_s": len(resps) / ((last - first) / 1e3)}
