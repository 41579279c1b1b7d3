"""CPU oracle for the sort / merge / scan path (numpy + ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, as the checker or the timed CPU
baseline, never as the product path.  See oracle.cpp for the reference
citations of every function.

Parity status of each function (pinned by tests/test_oracle.py):
  generate            pinned: SplitMix64 known-answer values + numpy restatement
  glibc_rand          pinned: glibc TYPE_3 additive-feedback known answer
  stable_sort         pinned: numpy stable argsort (independent implementation)
  segmented_stable_sort pinned: numpy lexsort by (segment, key)
  stable_sort_radix   pinned: numpy stable argsort of the bit keys; float
                      special values ordered by IEEE total order
  merge               pinned: numpy stable argsort of the concatenation
  scans / reduces     pinned: numpy cumsum in uint32 / exact int64 arithmetic
  merge_path          pinned: brute force count over the std::merge output
  mergesort_range     pinned: brute-force enumeration of tile coverage
  mergesort_schedule  pinned: equals stable_sort on many shapes
  radix_schedule      pinned: counts == bincount, offsets == brute force,
                      result == stable_sort_radix
  scan_schedule_u32   pinned: equals exclusive_scan_u32
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
KT = {"u32": 0, "i32": 1, "u64": 2, "f32": 3, "i64": 4, "f64": 5}
NP = {"u32": np.uint32, "i32": np.int32, "u64": np.uint64, "f32": np.float32, "i64": np.int64,
      "f64": np.float64}
WIDE = ("u64", "i64", "f64")
GEN = {"u32": 0, "lowentropy": 1, "i32_0_16": 2, "f32_01": 3, "u64": 4, "f32_pm1": 5, "iota": 6}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, c_int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.oracle_generate.argtypes = [vp, i64, c_int, c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_glibc_rand.argtypes = [i64, vp]
        L.oracle_stable_sort.argtypes = [c_int, vp, vp, i64, c_int]
        L.oracle_stable_sort_timed.argtypes = [c_int, vp, vp, i64, c_int]
        L.oracle_stable_sort_timed.restype = ctypes.c_double
        L.oracle_stable_sort_threads_timed.argtypes = [c_int, vp, vp, i64, c_int, c_int]
        L.oracle_stable_sort_threads_timed.restype = ctypes.c_double
        L.oracle_stable_sort_radix.argtypes = [c_int, vp, vp, i64]
        L.oracle_segmented_stable_sort.argtypes = [c_int, vp, vp, i64, vp, i64, c_int]
        L.oracle_std_sort_u32.argtypes = [vp, i64]
        L.oracle_segmented_scan_u32.argtypes = [vp, vp, i64, vp, i64, c_int, ctypes.c_uint32]
        L.oracle_segmented_scan_f64.argtypes = [vp, vp, i64, vp, i64, c_int, ctypes.c_double]
        L.oracle_merge.argtypes = [c_int, vp, vp, i64, vp, vp, i64, vp, vp, c_int]
        for n in ("oracle_exclusive_scan_u32", "oracle_inclusive_scan_u32"):
            getattr(L, n).argtypes = [vp, vp, i64, ctypes.c_uint32]
        for n in ("oracle_inclusive_scan_f64", "oracle_exclusive_scan_f64"):
            getattr(L, n).argtypes = [vp, vp, i64, ctypes.c_double]
        L.oracle_inclusive_scan_f32.argtypes = [vp, vp, i64]
        L.oracle_reduce_u32.argtypes = [vp, i64, ctypes.c_uint32]
        L.oracle_reduce_u32.restype = ctypes.c_uint32
        L.oracle_reduce_f64.argtypes = [vp, i64, ctypes.c_double]
        L.oracle_reduce_f64.restype = ctypes.c_double
        L.oracle_reduce_f32.argtypes = [vp, i64]
        L.oracle_reduce_f32.restype = ctypes.c_float
        L.oracle_merge_path.argtypes = [c_int, vp, i64, vp, i64, i64, c_int, c_int]
        L.oracle_merge_path.restype = i64
        L.oracle_mergesort_range.argtypes = [i64, i64, i64, i64, i64, i64, c_int, vp]
        L.oracle_mergesort_schedule.argtypes = [c_int, vp, vp, i64, c_int, c_int, c_int, c_int]
        L.oracle_radix_schedule.argtypes = [c_int, vp, vp, i64, i64, vp, vp]
        L.oracle_radix_bits.argtypes = [c_int, vp, vp, i64, c_int]
        L.oracle_scan_schedule_u32.argtypes = [vp, i64, i64, ctypes.c_uint32]
        L.oracle_set_plant.argtypes = [c_int]
        _lib = L
    return _lib


class planted:
    """Context manager that plants a known mistake in the oracle's
    restatements (1: ties to B in merge path; 2: mergesort range fix-up also on
    the last CTA of a coop group; 3: radix offsets in (tile, bin) order; 4: scan downsweep
    without carry-in), so a test can show its check rejects it."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().oracle_set_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_plant(0)
        return False


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def generate(count, kind, seed=0, offset=0):
    eb = 8 if kind == "u64" else 4
    dt = {"u32": np.uint32, "lowentropy": np.uint32, "i32_0_16": np.int32, "f32_01": np.float32,
          "u64": np.uint64, "f32_pm1": np.float32, "iota": np.uint32}[kind]
    out = np.empty(count, dtype=dt)
    lib().oracle_generate(_p(out), count, eb, GEN[kind], seed, offset)
    return out


def glibc_rand(n):
    out = np.empty(n, dtype=np.uint32)
    lib().oracle_glibc_rand(n, _p(out))
    return out


def stable_sort(keys, vals=None, kt="u32", greater=False):
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if v is not None and len(v) != len(k):
        raise ValueError("one value per key")
    lib().oracle_stable_sort(KT[kt], _p(k), _p(v), len(k), int(greater))
    return (k, v) if vals is not None else k


def f64_from_u64(x):
    """Finite doubles in [-1, 1) from 64-bit words (53 high bits), for key tests."""
    return (np.asarray(x, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0


def wide_keys(kt, x):
    """Keys of type kt from generator words x (u32 words for 4-byte types,
    u64 words for 8-byte types)."""
    if kt == "i32":
        return x.view(np.int32)
    if kt == "i64":
        return x.view(np.int64)
    if kt == "f64":
        return f64_from_u64(x)
    return x


def segmented_stable_sort(keys, segments, vals=None, kt="u32", greater=False):
    """std::stable_sort of every segment (segments = int32 starts)."""
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    sg = np.ascontiguousarray(segments, dtype=np.int32)
    lib().oracle_segmented_stable_sort(KT[kt], _p(k), _p(v), len(k), _p(sg), len(sg), int(greater))
    return (k, v) if vals is not None else k


def stable_sort_timed(keys, vals=None, kt="u32", greater=False):
    """Seconds std::stable_sort takes on a copy of (keys, vals) (the CPU
    baseline; copying and pair packing are not timed)."""
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if v is not None and len(v) != len(k):
        raise ValueError("one value per key")
    return float(lib().oracle_stable_sort_timed(KT[kt], _p(k), _p(v), len(k), int(greater)))


def stable_sort_threads(keys, vals=None, kt="u32", greater=False, threads=4):
    """The multi-threaded stable sort (one std::stable_sort per chunk, then
    merge rounds split by merge path): ((keys, vals) sorted, seconds)."""
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if v is not None and len(v) != len(k):
        raise ValueError("one value per key")
    dt = float(lib().oracle_stable_sort_threads_timed(KT[kt], _p(k), _p(v), len(k), int(greater),
                                                       int(threads)))
    return ((k, v) if vals is not None else k), dt


def stable_sort_radix(keys, vals=None, kt="u32"):
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if v is not None and len(v) != len(k):
        raise ValueError("one value per key")
    lib().oracle_stable_sort_radix(KT[kt], _p(k), _p(v), len(k))
    return (k, v) if vals is not None else k


def std_sort_u32(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    lib().oracle_std_sort_u32(_p(k), len(k))
    return k


def merge(a, b, a_vals=None, b_vals=None, kt="u32", greater=False):
    a = np.ascontiguousarray(a, dtype=NP[kt])
    b = np.ascontiguousarray(b, dtype=NP[kt])
    c = np.empty(len(a) + len(b), dtype=NP[kt])
    kv = a_vals is not None
    av = np.ascontiguousarray(a_vals, dtype=np.uint32) if kv else None
    bv = np.ascontiguousarray(b_vals, dtype=np.uint32) if kv else None
    if kv and (len(av) != len(a) or len(bv) != len(b)):
        raise ValueError("one value per key")
    cv = np.empty(len(c), dtype=np.uint32) if kv else None
    lib().oracle_merge(KT[kt], _p(a), _p(av), len(a), _p(b), _p(bv), len(b), _p(c), _p(cv),
                       int(greater))
    return (c, cv) if kv else c


def exclusive_scan_u32(x, init=0):
    x = np.ascontiguousarray(x).view(np.uint32)
    out = np.empty_like(x)
    lib().oracle_exclusive_scan_u32(_p(x), _p(out), len(x), init & 0xFFFFFFFF)
    return out


def inclusive_scan_u32(x, init=0):
    x = np.ascontiguousarray(x).view(np.uint32)
    out = np.empty_like(x)
    lib().oracle_inclusive_scan_u32(_p(x), _p(out), len(x), init & 0xFFFFFFFF)
    return out


def inclusive_scan_f64(x, init=0.0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(len(x), dtype=np.float64)
    lib().oracle_inclusive_scan_f64(_p(x), _p(out), len(x), init)
    return out


def exclusive_scan_f64(x, init=0.0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(len(x), dtype=np.float64)
    lib().oracle_exclusive_scan_f64(_p(x), _p(out), len(x), init)
    return out


def segmented_scan_u32(x, segments, inclusive=False, init=0):
    """Per-segment std::exclusive_scan / inclusive_scan mod 2^32 from init."""
    x = np.ascontiguousarray(x).view(np.uint32)
    sg = np.ascontiguousarray(segments, dtype=np.int32)
    out = np.empty_like(x)
    lib().oracle_segmented_scan_u32(_p(x), _p(out), len(x), _p(sg), len(sg), int(inclusive),
                                    init & 0xFFFFFFFF)
    return out


def segmented_scan_f64(x, segments, inclusive=False, init=0.0):
    """Per-segment float64 prefix sums of float32 data from init."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    sg = np.ascontiguousarray(segments, dtype=np.int32)
    out = np.empty(len(x), dtype=np.float64)
    lib().oracle_segmented_scan_f64(_p(x), _p(out), len(x), _p(sg), len(sg), int(inclusive),
                                    float(init))
    return out


def inclusive_scan_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().oracle_inclusive_scan_f32(_p(x), _p(out), len(x))
    return out


def reduce_u32(x, init=0):
    x = np.ascontiguousarray(x).view(np.uint32)
    return int(lib().oracle_reduce_u32(_p(x), len(x), init & 0xFFFFFFFF))


def reduce_f64(x, init=0.0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().oracle_reduce_f64(_p(x), len(x), init))


def reduce_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().oracle_reduce_f32(_p(x), len(x)))


def merge_path(a, b, diag, kt="u32", greater=False, upper=False):
    a = np.ascontiguousarray(a, dtype=NP[kt])
    b = np.ascontiguousarray(b, dtype=NP[kt])
    return int(lib().oracle_merge_path(KT[kt], _p(a), len(a), _p(b), len(b), diag, int(greater),
                                       int(upper)))


def mergesort_range(count, partition, coop, spacing, mp0=None, mp1=None):
    out = np.zeros(4, dtype=np.int64)
    use = mp0 is not None
    lib().oracle_mergesort_range(count, partition, coop, spacing, mp0 or 0, mp1 or 0, int(use),
                                 _p(out))
    return tuple(int(x) for x in out)


def mergesort_schedule(keys, vals=None, nt=128, vt=7, kt="u32", greater=False,
                       sort_indices=False):
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None
    if vals is not None or sort_indices:
        v = (np.zeros(len(k), np.uint32) if vals is None
             else np.ascontiguousarray(vals, dtype=np.uint32).copy())
    lib().oracle_mergesort_schedule(KT[kt], _p(k), _p(v), len(k), nt, vt, int(greater),
                                    int(sort_indices))
    return (k, v) if v is not None else k


def radix_schedule(keys, vals=None, nv=4096, kt="u32", want_first_pass=False):
    k = np.ascontiguousarray(keys, dtype=NP[kt]).copy()
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32).copy()
    tiles = (len(k) + nv - 1) // nv
    counts = np.zeros(tiles * 256, np.uint32) if want_first_pass else None
    offs = np.zeros(tiles * 256, np.uint64) if want_first_pass else None
    lib().oracle_radix_schedule(KT[kt], _p(k), _p(v), len(k), nv, _p(counts), _p(offs))
    out = (k, v) if vals is not None else k
    if want_first_pass:
        return out, counts.reshape(tiles, 256), offs.reshape(tiles, 256)
    return out


def radix_bits(keys, kt="f32", inverse=False):
    if inverse:
        src = np.ascontiguousarray(keys, dtype=np.uint64 if kt in WIDE else np.uint32)
        out = np.empty(len(src), dtype=NP[kt])
    else:
        src = np.ascontiguousarray(keys, dtype=NP[kt])
        out = np.empty(len(src), dtype=np.uint64 if kt in WIDE else np.uint32)
    lib().oracle_radix_bits(KT[kt], _p(src), _p(out), len(src), int(inverse))
    return out


def scan_schedule_u32(x, nv=896, init=0):
    d = np.ascontiguousarray(x).view(np.uint32).copy()
    lib().oracle_scan_schedule_u32(_p(d), len(d), nv, init & 0xFFFFFFFF)
    return d
