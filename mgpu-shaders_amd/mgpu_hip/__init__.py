"""Python mirror of the mgpu sort / merge / scan call surface over libmgpu_hip.so.

The names and argument meanings follow the reference's `inc/mgpu/vk` host
templates (mergesort_keys / mergesort_kv / mergesort_indices,
inc/mgpu/vk/mergesort.hxx:10-103; merge, inc/mgpu/vk/merge.hxx:12-59;
radix_sort, inc/mgpu/vk/radix.hxx:195-198; scan, inc/mgpu/vk/scan.hxx:10-13).
Sorts and scans are in place like the reference; merge writes its output.
Every call goes through the C ABI in include/mgpu_hip.h: a sizing call with a
NULL aux buffer, then the real call with an aux buffer of that size.  There is
no CPU fallback: without the built library every entry point raises.

Tensors are device memory (torch is only used for allocation and streams).
32-bit unsigned keys may be passed as torch.uint32 or as torch.int32 with
key_type="u32"; 64-bit unsigned keys as torch.uint64 or torch.int64 with
key_type="u64".
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgpu_hip.so")

MGPU_OK, MGPU_INVALID_ARG, MGPU_AUX_TOO_SMALL, MGPU_HIP_ERROR, MGPU_UNSUPPORTED = range(5)
LESS, GREATER = 0, 1
EXCLUSIVE, INCLUSIVE = 0, 1
GEN_KINDS = {"u32": 0, "lowentropy": 1, "i32_0_16": 2, "f32_01": 3, "u64": 4, "f32_pm1": 5,
             "iota": 6, "sorted_u32": 7, "sorted_u64": 8, "sorted_f32": 9}
KEY_TYPES = ("u32", "i32", "u64", "f32", "i64", "f64")
_CTYPE = {"u32": ctypes.c_uint32, "i32": ctypes.c_int32, "u64": ctypes.c_uint64,
          "f32": ctypes.c_float, "i64": ctypes.c_int64, "f64": ctypes.c_double}
_DTYPE_KEY = {torch.int32: "i32", torch.float32: "f32", torch.uint32: "u32",
              torch.uint64: "u64", torch.int64: "i64", torch.float64: "f64"}
_KEY_BYTES = {"u32": 4, "i32": 4, "u64": 8, "f32": 4, "i64": 8, "f64": 8}


class MgpuError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        lib = _lib_or_none()
        name = lib.mgpu_status_string(status).decode() if lib else str(status)
        extra = ""
        if status == MGPU_HIP_ERROR and lib:
            extra = " (hipError_t %d)" % lib.mgpu_last_hip_error()
        super().__init__("%s failed: %s%s" % (where, name, extra))


_lib = None


def _lib_or_none():
    return _lib


def load():
    """Load libmgpu_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmgpu_hip.so not built (run python mgpu-shaders_amd/build.py): "
                          + LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, i64, c_int = ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int64, ctypes.c_int
    lib.mgpu_status_string.restype = ctypes.c_char_p
    lib.mgpu_status_string.argtypes = [c_int]
    lib.mgpu_last_hip_error.restype = c_int
    lib.mgpu_version.restype = c_int
    for k in KEY_TYPES:
        for name, args in [
            ("mgpu_mergesort_keys_%s", [vp, sz, vp, vp, i64, c_int]),
            ("mgpu_mergesort_kv_%s_u32", [vp, sz, vp, vp, vp, i64, c_int]),
            ("mgpu_mergesort_indices_%s", [vp, sz, vp, vp, vp, i64, c_int]),
            ("mgpu_merge_keys_%s", [vp, sz, vp, vp, i64, vp, i64, vp, c_int]),
            ("mgpu_merge_kv_%s_u32", [vp, sz, vp, vp, vp, i64, vp, vp, i64, vp, vp, c_int]),
            ("mgpu_radix_sort_keys_%s", [vp, sz, vp, vp, i64]),
            ("mgpu_merge_path_partitions_%s", [vp, vp, i64, vp, i64, vp, i64, c_int, c_int]),
            ("mgpu_radix_sort_kv_%s_u32", [vp, sz, vp, vp, vp, i64]),
        ]:
            f = getattr(lib, name % k)
            f.restype = c_int
            f.argtypes = args
    for t in ("i32", "u32", "f32", "u64"):
        ct = _CTYPE[t]
        f = getattr(lib, "mgpu_scan_exclusive_%s" % t)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, i64, ct]
        f = getattr(lib, "mgpu_scan_%s" % t)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, vp, i64, c_int, ct]
        f = getattr(lib, "mgpu_reduce_%s" % t)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, i64, vp, ct]
    for t in ("i32", "u32", "f32"):
        f = getattr(lib, "mgpu_segmented_scan_%s" % t)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, vp, i64, vp, i64, c_int, _CTYPE[t]]
    class SegsortTuning(ctypes.Structure):
        _fields_ = [("num_buckets", ctypes.c_int), ("cap", ctypes.c_int * SEGSORT_MAX_BUCKETS),
                    ("shape", ctypes.c_int * SEGSORT_MAX_BUCKETS)]
    global _SegsortTuning
    _SegsortTuning = SegsortTuning
    tp = ctypes.POINTER(SegsortTuning)
    for k in KEY_TYPES:
        f = getattr(lib, "mgpu_segmented_sort_keys_%s" % k)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, i64, vp, i64, c_int, tp]
        f = getattr(lib, "mgpu_segmented_sort_kv_%s_u32" % k)
        f.restype, f.argtypes = c_int, [vp, sz, vp, vp, vp, i64, vp, i64, c_int, tp]
        f = getattr(lib, "mgpu_segmented_sort_autotune_%s" % k)
        f.restype, f.argtypes = c_int, [vp, c_int, c_int, tp, ctypes.POINTER(ctypes.c_double)]
    lib.mgpu_segsort_shapes.restype = c_int
    lib.mgpu_segsort_shapes.argtypes = [ctypes.POINTER(c_int), ctypes.POINTER(c_int), c_int]
    lib.mgpu_capture_begin.restype = c_int
    lib.mgpu_capture_begin.argtypes = [vp]
    lib.mgpu_capture_end.restype = c_int
    lib.mgpu_capture_end.argtypes = [vp, ctypes.POINTER(vp)]
    lib.mgpu_graph_launch.restype = c_int
    lib.mgpu_graph_launch.argtypes = [vp, vp]
    lib.mgpu_graph_destroy.restype = None
    lib.mgpu_graph_destroy.argtypes = [vp]
    lib.mgpu_generate.restype = c_int
    lib.mgpu_generate.argtypes = [vp, vp, i64, c_int, c_int, ctypes.c_uint64, ctypes.c_uint64]
    lib.mgpu_copy.argtypes = [vp, vp, vp, i64]
    lib.mgpu_set_radix_algorithm.restype = None
    lib.mgpu_set_radix_algorithm.argtypes = [c_int]
    lib.mgpu_set_mergesort_passes.restype = None
    lib.mgpu_set_mergesort_passes.argtypes = [c_int]
    lib.mgpu_mergesort_passes.restype = c_int
    lib.mgpu_mergesort_passes.argtypes = []
    lib.mgpu_radix_algorithm.restype = c_int
    lib.mgpu_set_lookback_spin_limit.restype = None
    lib.mgpu_set_lookback_spin_limit.argtypes = [ctypes.c_uint32]
    lib.mgpu_lookback_spin_limit.restype = ctypes.c_uint32
    lib.mgpu_lookback_fallbacks.restype = ctypes.c_int64
    lib.mgpu_lookback_fallbacks.argtypes = [c_int]
    lib.mgpu_timing_enable.restype = None
    lib.mgpu_timing_enable.argtypes = [c_int]
    lib.mgpu_timing_collect.restype = c_int
    lib.mgpu_timing_collect.argtypes = [ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_double), c_int]
    _lib = lib
    return lib


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _check(status, where):
    if status != MGPU_OK:
        raise MgpuError(status, where)


def _key_type(t, key_type):
    if key_type is not None:
        if key_type not in KEY_TYPES:
            raise ValueError("key_type must be one of %s" % (KEY_TYPES,))
        if t.element_size() != _KEY_BYTES[key_type]:
            raise ValueError("tensor element size does not match key_type %s" % key_type)
        return key_type
    if t.dtype not in _DTYPE_KEY:
        raise ValueError("cannot infer key type of %s; pass key_type" % t.dtype)
    return _DTYPE_KEY[t.dtype]


def _comp(comp):
    if comp in ("less", LESS):
        return LESS
    if comp in ("greater", GREATER):
        return GREATER
    raise ValueError("comp must be 'less' or 'greater'")


def _device_check(*ts):
    for t in ts:
        if t is not None:
            if not t.is_cuda:
                raise ValueError("mgpu_hip operates on device tensors")
            if not t.is_contiguous():
                raise ValueError("mgpu_hip needs contiguous tensors")


def _run_two_phase(fn, args, where, aux=None, stream=None):
    """Sizing call (aux=NULL) then the real call; allocates aux unless given.

    The call is enqueued on `stream` (default: torch's current stream).  An
    aux tensor is allocated on that stream, and when `stream` is not torch's
    current stream the aux is also marked as used by it (record_stream), so
    torch's caching allocator does not hand its memory to another allocation
    while the library may still be writing it, whether or not the caller
    keeps the returned aux."""
    lib = load()
    size = ctypes.c_size_t(0)
    _check(fn(None, ctypes.byref(size), _stream(stream), *args), where + " (sizing)")
    if aux is None:
        if stream is not None:
            with torch.cuda.stream(stream):
                aux = torch.empty(max(int(size.value), 1), dtype=torch.uint8, device="cuda")
        else:
            aux = torch.empty(max(int(size.value), 1), dtype=torch.uint8, device="cuda")
    if stream is not None and stream != torch.cuda.current_stream():
        aux.record_stream(stream)
    cap = ctypes.c_size_t(aux.numel() * aux.element_size())
    _check(fn(_ptr(aux), ctypes.byref(cap), _stream(stream), *args), where)
    return aux


def aux_size(fn_name, *args, stream=None):
    """Bytes of aux the named C entry point needs for these arguments."""
    lib = load()
    size = ctypes.c_size_t(0)
    _check(getattr(lib, fn_name)(None, ctypes.byref(size), _stream(stream), *args), fn_name)
    return int(size.value)


# ---------------------------------------------------------------------------
# Mergesort (inc/mgpu/vk/mergesort.hxx:10-103)

def mergesort_keys(keys, comp="less", key_type=None, aux=None, stream=None):
    _device_check(keys)
    k = _key_type(keys, key_type)
    fn = getattr(load(), "mgpu_mergesort_keys_%s" % k)
    return _run_two_phase(fn, [_ptr(keys), keys.numel(), _comp(comp)], "mergesort_keys", aux, stream)


def mergesort_kv(keys, vals, comp="less", key_type=None, aux=None, stream=None):
    _device_check(keys, vals)
    if vals.numel() != keys.numel() or vals.element_size() != 4:
        raise ValueError("vals must be 32-bit with one per key")
    k = _key_type(keys, key_type)
    fn = getattr(load(), "mgpu_mergesort_kv_%s_u32" % k)
    return _run_two_phase(fn, [_ptr(keys), _ptr(vals), keys.numel(), _comp(comp)],
                          "mergesort_kv", aux, stream)


def mergesort_indices(keys, indices, comp="less", key_type=None, aux=None, stream=None):
    _device_check(keys, indices)
    if indices.numel() != keys.numel() or indices.dtype != torch.int32:
        raise ValueError("indices must be int32 with one per key")
    k = _key_type(keys, key_type)
    fn = getattr(load(), "mgpu_mergesort_indices_%s" % k)
    return _run_two_phase(fn, [_ptr(keys), _ptr(indices), keys.numel(), _comp(comp)],
                          "mergesort_indices", aux, stream)


# ---------------------------------------------------------------------------
# Merge (inc/mgpu/vk/merge.hxx:12-59)

def _check_merge_shapes(a_keys, b_keys, c_keys, a_vals=None, b_vals=None, c_vals=None):
    """The C entry points take raw pointers and counts from a and b: c and the
    value arrays must hold exactly what those counts imply."""
    if c_keys.numel() != a_keys.numel() + b_keys.numel():
        raise ValueError("c must hold a + b elements")
    if not a_keys.dtype == b_keys.dtype == c_keys.dtype:
        raise ValueError("a, b and c keys must have one dtype")
    for keys, vals, name in ((a_keys, a_vals, "a"), (b_keys, b_vals, "b"), (c_keys, c_vals, "c")):
        if vals is not None and (vals.numel() != keys.numel() or vals.element_size() != 4):
            raise ValueError("%s_vals must be 32-bit with one per %s key" % (name, name))


def merge_keys(a_keys, b_keys, c_keys, comp, key_type=None, aux=None, stream=None):
    _check_merge_shapes(a_keys, b_keys, c_keys)
    _device_check(a_keys, b_keys, c_keys)
    k = _key_type(a_keys, key_type)
    fn = getattr(load(), "mgpu_merge_keys_%s" % k)
    return _run_two_phase(fn, [_ptr(a_keys), a_keys.numel(), _ptr(b_keys), b_keys.numel(),
                               _ptr(c_keys), _comp(comp)], "merge_keys", aux, stream)


def merge(a_keys, a_vals, b_keys, b_vals, c_keys, c_vals, comp, key_type=None, aux=None,
          stream=None):
    _check_merge_shapes(a_keys, b_keys, c_keys, a_vals, b_vals, c_vals)
    _device_check(a_keys, a_vals, b_keys, b_vals, c_keys, c_vals)
    k = _key_type(a_keys, key_type)
    fn = getattr(load(), "mgpu_merge_kv_%s_u32" % k)
    return _run_two_phase(fn, [_ptr(a_keys), _ptr(a_vals), a_keys.numel(), _ptr(b_keys),
                               _ptr(b_vals), b_keys.numel(), _ptr(c_keys), _ptr(c_vals),
                               _comp(comp)], "merge", aux, stream)


def num_merge_partitions(count, spacing):
    """Entries of a partition array (inc/mgpu/common/cta_merge.hxx:6-8)."""
    return (count + spacing - 1) // spacing + 1


def merge_path_partitions(a_keys, b_keys, spacing, comp="less", bounds="lower", mp=None,
                          key_type=None, stream=None):
    """vk::merge_path_partitions<bounds> (inc/mgpu/vk/partition.hxx:9-20):
    mp[i] = A elements among the first min(i * spacing, |a| + |b|) outputs of
    the merge (ties take A first with bounds="lower", B first with "upper").
    Returns mp, an int64 device tensor of num_merge_partitions entries."""
    _device_check(a_keys, b_keys)
    if a_keys.dtype != b_keys.dtype:
        raise ValueError("a and b keys must have one dtype")
    if spacing <= 0:
        raise ValueError("spacing must be positive")
    if bounds not in ("lower", "upper"):
        raise ValueError("bounds must be 'lower' or 'upper'")
    k = _key_type(a_keys, key_type)
    n = num_merge_partitions(a_keys.numel() + b_keys.numel(), spacing)
    if mp is None:
        if stream is not None:
            with torch.cuda.stream(stream):
                mp = torch.empty(n, dtype=torch.int64, device="cuda")
        else:
            mp = torch.empty(n, dtype=torch.int64, device="cuda")
    elif mp.dtype != torch.int64 or mp.numel() < n:
        raise ValueError("mp must be int64 with at least %d entries" % n)
    _device_check(mp)
    fn = getattr(load(), "mgpu_merge_path_partitions_%s" % k)
    _check(fn(_stream(stream), _ptr(a_keys), a_keys.numel(), _ptr(b_keys), b_keys.numel(),
              _ptr(mp), spacing, _comp(comp), 0 if bounds == "lower" else 1),
           "merge_path_partitions")
    return mp


# ---------------------------------------------------------------------------
# Radix sort (inc/mgpu/vk/radix.hxx:195-470, plus values)

def radix_sort(keys, vals=None, key_type=None, aux=None, stream=None):
    _device_check(keys, vals)
    k = _key_type(keys, key_type)
    lib = load()
    if vals is None:
        fn = getattr(lib, "mgpu_radix_sort_keys_%s" % k)
        return _run_two_phase(fn, [_ptr(keys), keys.numel()], "radix_sort", aux, stream)
    if vals.numel() != keys.numel() or vals.element_size() != 4:
        raise ValueError("vals must be 32-bit with one per key")
    fn = getattr(lib, "mgpu_radix_sort_kv_%s_u32" % k)
    return _run_two_phase(fn, [_ptr(keys), _ptr(vals), keys.numel()], "radix_sort", aux, stream)


# ---------------------------------------------------------------------------
# Scan / reduce (inc/mgpu/vk/scan.hxx:10-115)

_SCAN_T = {torch.int32: "i32", torch.uint32: "u32", torch.float32: "f32", torch.uint64: "u64"}
_SCAN_BYTES = {"i32": 4, "u32": 4, "f32": 4, "u64": 8}


def _scan_type(t, elem_type, out=None):
    """Scan element type of t (or elem_type), checked against t's element size
    and, when given, the output's: the entry points write sizeof(type) bytes
    per element, so a narrower output would be overrun."""
    if elem_type is not None:
        if elem_type not in _SCAN_BYTES:
            raise ValueError("elem_type must be one of %s" % (tuple(_SCAN_BYTES),))
        t_type = elem_type
    elif t.dtype in _SCAN_T:
        t_type = _SCAN_T[t.dtype]
    else:
        raise ValueError("scan supports int32, uint32, float32 and uint64")
    if t.element_size() != _SCAN_BYTES[t_type]:
        raise ValueError("input element size does not match %s" % t_type)
    if out is not None and out.element_size() != t.element_size():
        raise ValueError("output element size must match the input's")
    return t_type


def scan(data, init=0, elem_type=None, aux=None, stream=None):
    """In-place exclusive plus-scan: data[i] = init + sum(data[:i])."""
    t = _scan_type(data, elem_type)
    _device_check(data)
    fn = getattr(load(), "mgpu_scan_exclusive_%s" % t)
    return _run_two_phase(fn, [_ptr(data), data.numel(), _CTYPE[t](init)], "scan", aux, stream)


def scan_into(input, output, inclusive=False, init=0, elem_type=None, aux=None, stream=None):
    """Out-of-place (or in-place when output is input) exclusive/inclusive scan."""
    if output.numel() != input.numel():
        raise ValueError("output must match input")
    t = _scan_type(input, elem_type, output)
    _device_check(input, output)
    fn = getattr(load(), "mgpu_scan_%s" % t)
    return _run_two_phase(fn, [_ptr(input), _ptr(output), input.numel(),
                               INCLUSIVE if inclusive else EXCLUSIVE, _CTYPE[t](init)],
                          "scan", aux, stream)


def segmented_scan(input, output, segments, inclusive=False, init=0, elem_type=None, aux=None,
                   stream=None):
    """One scan per segment, each from init; segments = int32 device tensor of
    ascending segment starts (element 0 always starts a segment)."""
    if output.numel() != input.numel():
        raise ValueError("output must match input")
    if segments.dtype != torch.int32:
        raise ValueError("segments must be int32")
    t = _scan_type(input, elem_type, output)
    _device_check(input, output, segments)
    if t == "u64":
        raise ValueError("segmented scan supports int32, uint32 and float32")
    fn = getattr(load(), "mgpu_segmented_scan_%s" % t)
    return _run_two_phase(fn, [_ptr(input), _ptr(output), input.numel(), _ptr(segments),
                               segments.numel(), INCLUSIVE if inclusive else EXCLUSIVE,
                               _CTYPE[t](init)], "segmented_scan", aux, stream)


def reduce(input, out, init=0, elem_type=None, aux=None, stream=None):
    """out[0] = init + sum(input) (out is a 1-element device tensor)."""
    if out.numel() < 1:
        raise ValueError("out must hold one element")
    t = _scan_type(input, elem_type, out)
    _device_check(input, out)
    fn = getattr(load(), "mgpu_reduce_%s" % t)
    return _run_two_phase(fn, [_ptr(input), input.numel(), _ptr(out), _CTYPE[t](init)],
                          "reduce", aux, stream)


# ---------------------------------------------------------------------------
# Segmented sort (batched; segments[i] = first element of segment i)

_SegsortTuning = None
SEGSORT_MAX_BUCKETS = 13  # MGPU_SEGSORT_MAX_BUCKETS of mgpu_hip.h


def segsort_shapes():
    """[(nt, vt)] of the blocksort shapes a bucket can use."""
    lib = load()
    nt = (ctypes.c_int * 64)()
    vt = (ctypes.c_int * 64)()
    n = lib.mgpu_segsort_shapes(nt, vt, 64)
    return [(nt[i], vt[i]) for i in range(min(n, 64))]


def _tuning_struct(tuning):
    if tuning is None:
        return None
    if len(tuning["cap"]) > SEGSORT_MAX_BUCKETS or len(tuning["shape"]) != len(tuning["cap"]):
        raise ValueError("a tuning table holds at most %d buckets, one shape per cap"
                         % SEGSORT_MAX_BUCKETS)
    t = _SegsortTuning()
    t.num_buckets = len(tuning["cap"])
    for b, (c, sh) in enumerate(zip(tuning["cap"], tuning["shape"])):
        t.cap[b] = c
        t.shape[b] = sh
    return ctypes.byref(t)


def segmented_sort(keys, segments, vals=None, comp="less", key_type=None, tuning=None, aux=None,
                   stream=None):
    """Stable sort of every segment of keys (+ vals) in place; segments is an
    int32 device tensor of segment starts."""
    _device_check(keys, vals, segments)
    if segments.dtype != torch.int32:
        raise ValueError("segments must be int32")
    k = _key_type(keys, key_type)
    lib = load()
    t = _tuning_struct(tuning)
    if vals is None:
        fn = getattr(lib, "mgpu_segmented_sort_keys_%s" % k)
        args = [_ptr(keys), keys.numel(), _ptr(segments), segments.numel(), _comp(comp), t]
    else:
        if vals.numel() != keys.numel() or vals.element_size() != 4:
            raise ValueError("vals must be 32-bit with one per key")
        fn = getattr(lib, "mgpu_segmented_sort_kv_%s_u32" % k)
        args = [_ptr(keys), _ptr(vals), keys.numel(), _ptr(segments), segments.numel(),
                _comp(comp), t]
    return _run_two_phase(fn, args, "segmented_sort", aux, stream)


def segmented_sort_autotune(key_type="u32", has_values=True, comp="less", stream=None):
    """Choose the blocksort shape of every size bucket on this device, by
    timing whole segmented sorts of a synthetic layout (coordinate descent
    over the buckets); the chosen table becomes the default for this
    key/value/comparator and is returned as {"cap": [...], "shape": [...],
    "ms": [[...]]}, ms[b][i] = the call's time with shape i in bucket b (-1
    where not tried)."""
    lib = load()
    t = _SegsortTuning()
    nshapes = len(segsort_shapes())
    ms = (ctypes.c_double * (SEGSORT_MAX_BUCKETS * nshapes))()
    _check(getattr(lib, "mgpu_segmented_sort_autotune_%s" % key_type)(
        _stream(stream), 1 if has_values else 0, _comp(comp), ctypes.byref(t), ms),
        "segmented_sort_autotune")
    nb = t.num_buckets
    return {"cap": [t.cap[b] for b in range(nb)], "shape": [t.shape[b] for b in range(nb)],
            "ms": [[ms[b * nshapes + i] for i in range(nshapes)] for b in range(nb)]}


# ---------------------------------------------------------------------------
# Graph capture of library calls

class Graph:
    """Record library calls made on `stream` inside `with Graph(stream) as g:`
    and replay them with g.launch()."""

    def __init__(self, stream):
        self.stream = stream
        self.handle = ctypes.c_void_p(0)

    def __enter__(self):
        _check(load().mgpu_capture_begin(_stream(self.stream)), "capture_begin")
        return self

    def __exit__(self, *exc):
        st = load().mgpu_capture_end(_stream(self.stream), ctypes.byref(self.handle))
        if exc[0] is None:
            _check(st, "capture_end")
        return False

    def launch(self, stream=None):
        _check(load().mgpu_graph_launch(self.handle, _stream(stream or self.stream)),
               "graph_launch")

    def __del__(self):
        if self.handle and _lib is not None:
            _lib.mgpu_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)


# ---------------------------------------------------------------------------
# Synthetic inputs and kernel timing

def generate(out, kind, seed=0, offset=0, stream=None):
    """Fill `out` with element (offset + i) of a SplitMix64 generator kind."""
    _device_check(out)
    _check(load().mgpu_generate(_stream(stream), _ptr(out), out.numel(), out.element_size(),
                                GEN_KINDS[kind], seed, offset), "generate")
    return out


def copy(dst, src, stream=None):
    """dst := src (contiguous device tensors of the same byte size, 16-byte
    aligned) with the library's plain HBM copy kernel."""
    _device_check(dst, src)
    nbytes = src.numel() * src.element_size()
    if not (dst.is_contiguous() and src.is_contiguous()) or \
            dst.numel() * dst.element_size() != nbytes:
        raise ValueError("copy needs contiguous tensors of the same byte size")
    _check(load().mgpu_copy(_stream(stream), _ptr(dst), _ptr(src), nbytes), "copy")
    return dst


RADIX_AUTO, RADIX_COUNT_SCAN_SCATTER, RADIX_SINGLE_PASS = 0, 1, 2


def set_radix_algorithm(algorithm):
    """0 = automatic, 1 = count + scan + scatter per pass, 2 = single pass."""
    load().mgpu_set_radix_algorithm(int(algorithm))


def radix_algorithm():
    return int(load().mgpu_radix_algorithm())


MERGESORT_AUTO, MERGESORT_TWO_WAY, MERGESORT_FOUR_WAY = 0, 2, 4


def set_mergesort_passes(ways):
    """0 = automatic, 2 = two-way merge passes only, 4 = four-way passes."""
    load().mgpu_set_mergesort_passes(int(ways))


def mergesort_passes():
    return int(load().mgpu_mergesort_passes())


def set_lookback_spin_limit(limit):
    """Polls a look-back waits for a predecessor before computing its part itself."""
    load().mgpu_set_lookback_spin_limit(int(limit))


def lookback_spin_limit():
    return int(load().mgpu_lookback_spin_limit())


def lookback_fallbacks(reset=False):
    """How often a look-back computed a predecessor's contribution (syncs)."""
    return int(load().mgpu_lookback_fallbacks(1 if reset else 0))


def timing_enable(enable=True):
    load().mgpu_timing_enable(1 if enable else 0)


def timing_collect():
    """{kernel_name: (launches, total_ms)} of launches since timing_enable."""
    lib = load()
    n = lib.mgpu_timing_collect(None, None, None, 0)
    names = (ctypes.c_char_p * max(n, 1))()
    launches = (ctypes.c_int * max(n, 1))()
    total = (ctypes.c_double * max(n, 1))()
    lib.mgpu_timing_collect(names, launches, total, n)
    return {names[i].decode(): (launches[i], total[i]) for i in range(n)}
