"""C ABI checks that need no GPU: exports, sizing calls, argument errors."""
import ctypes
import os
import re
import subprocess

import pytest

import mgpu_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mgpu_hip.h")


def declared_functions():
    pre = subprocess.run(["gcc", "-E", "-P", HEADER], capture_output=True, text=True, check=True)
    return sorted(set(re.findall(r"\b(mgpu_[a-z0-9_]+)\s*\(", pre.stdout)))


def test_library_exports_every_declared_symbol():
    lib = mgpu_hip.load()
    names = declared_functions()
    assert len(names) >= 40
    nm = subprocess.run(["nm", "-D", "--defined-only", mgpu_hip.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r" T (mgpu_\w+)", nm))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None


def test_status_strings():
    lib = mgpu_hip.load()
    assert lib.mgpu_status_string(0) == b"MGPU_OK"
    assert lib.mgpu_status_string(2) == b"MGPU_AUX_TOO_SMALL"
    assert lib.mgpu_version() >= 100


def test_radix_algorithm_setting():
    """0 automatic, 1 count + scan + scatter, 2 single pass; anything else
    falls back to automatic.  Sizing does not depend on the setting."""
    old = mgpu_hip.radix_algorithm()
    one = ctypes.c_void_p(1)
    try:
        sizes = []
        for a in (0, 1, 2):
            mgpu_hip.set_radix_algorithm(a)
            assert mgpu_hip.radix_algorithm() == a
            sizes.append(_size(mgpu_hip.load().mgpu_radix_sort_kv_u32_u32, one, one, 1 << 20))
        assert sizes[0] == sizes[1] == sizes[2]
        for bad in (-1, 3, 99):
            mgpu_hip.set_radix_algorithm(bad)
            assert mgpu_hip.radix_algorithm() == 0
    finally:
        mgpu_hip.set_radix_algorithm(old)


def test_mergesort_passes_setting():
    """0 automatic, 2 two-way, 4 four-way; anything else falls back to
    automatic.  Sizing does not depend on the setting."""
    old = mgpu_hip.mergesort_passes()
    one = ctypes.c_void_p(1)
    try:
        sizes = []
        for w in (0, 2, 4):
            mgpu_hip.set_mergesort_passes(w)
            assert mgpu_hip.mergesort_passes() == w
            sizes.append(_size(mgpu_hip.load().mgpu_mergesort_kv_u32_u32, one, one, 1 << 26, 0))
        assert sizes[0] == sizes[1] == sizes[2]
        for bad in (-1, 1, 3, 8):
            mgpu_hip.set_mergesort_passes(bad)
            assert mgpu_hip.mergesort_passes() == 0
    finally:
        mgpu_hip.set_mergesort_passes(old)


def _size(fn, *args):
    size = ctypes.c_size_t(0)
    st = fn(None, ctypes.byref(size), None, *args)
    return st, size.value


def r256(x):
    return (x + 255) // 256 * 256


def merge4_samples(n, nv):
    """Samples of the largest four-way pass: one per tile of X and of Y per
    group of four runs (levels taken two at a time, after a two-way level
    when their number is odd)."""
    levels = max(0, (-(-n // nv) - 1).bit_length())
    most = 0
    for level in range(levels & 1, levels, 2):
        run = nv << level
        most = max(most, 2 * (-(-n // (4 * run))) * (2 * run // nv))
    return most


def test_sizing_mergesort_matches_formula():
    lib = mgpu_hip.load()
    n = 1 << 20
    nv = 512 * 19  # mergesort tile of 32-bit keys
    tiles = -(-n // nv)
    assert merge4_samples(n, nv) == 128
    parts = r256(16 * (tiles + 1)) + r256(8 * merge4_samples(n, nv))
    st, sz = _size(lib.mgpu_mergesort_kv_u32_u32, ctypes.c_void_p(1), ctypes.c_void_p(1), n, 0)
    assert st == 0
    assert sz == parts + r256(4 * n) + r256(4 * n)
    st, sz1 = _size(lib.mgpu_mergesort_keys_u32, ctypes.c_void_p(1), n, 0)
    assert sz1 == parts + r256(4 * n)
    # One tile: no aux bytes (vk/mergesort.hxx:22-23).
    st, sz = _size(lib.mgpu_mergesort_keys_u32, ctypes.c_void_p(1), 100, 0)
    assert st == 0 and sz == 0


def test_sizing_accumulates():
    lib = mgpu_hip.load()
    size = ctypes.c_size_t(1000)
    st = lib.mgpu_merge_keys_u32(None, ctypes.byref(size), None, ctypes.c_void_p(1), 5000,
                                 ctypes.c_void_p(1), 5000, ctypes.c_void_p(1), 0)
    assert st == 0
    assert size.value == 1000 + r256(8 * (-(-10000 // (512 * 23)) + 1))  # merge tile


def radix_aux(n, key_bytes, val_bytes, nv, passes):
    """Ping-pong buffers + max(count + scan path, single-pass path)."""
    tiles = -(-n // nv)
    scan_aux = r256(8 * (-(-tiles * 256 // 4096)))  # counts scan: 4096-element tiles
    # Count path: per-tile counts, their scan, and for 64-bit keys one digit
    # byte per key.
    scan_path = r256(4 * tiles * 256) + scan_aux + (r256(n) if key_bytes == 8 else 0)
    # Single pass: [passes][tiles][256] status words, all-pass counts, bases.
    single_pass = r256(4 * tiles * 256 * passes) + 2 * r256(4 * passes * 256)
    return r256(key_bytes * n) + (r256(val_bytes * n) if val_bytes else 0) + \
        max(scan_path, single_pass)


def test_sizing_radix_and_scan():
    lib = mgpu_hip.load()
    one = ctypes.c_void_p(1)
    # Tile shape by size: 32-bit kv 512 x 8 up to 6 * 2^20 keys, 512 x 20 above;
    # 32-bit keys 512 x 8 / 512 x 12 / 512 x 24 (cuts at 2^22, 2^25); 64-bit
    # keys 512 x 8 up to 2^22, 512 x 12 above.
    for fn, args, n, kb, vb, nv, passes in [
            (lib.mgpu_radix_sort_kv_u32_u32, (one, one), 1 << 20, 4, 4, 4096, 4),
            (lib.mgpu_radix_sort_kv_u32_u32, (one, one), 7 << 20, 4, 4, 10240, 4),
            (lib.mgpu_radix_sort_keys_f32, (one,), 1 << 22, 4, 0, 4096, 4),
            (lib.mgpu_radix_sort_keys_f32, (one,), (1 << 22) + 1, 4, 0, 6144, 4),
            (lib.mgpu_radix_sort_keys_u32, (one,), (1 << 25) + 1, 4, 0, 12288, 4),
            (lib.mgpu_radix_sort_kv_u64_u32, (one, one), 1 << 22, 8, 4, 4096, 8),
            (lib.mgpu_radix_sort_kv_u64_u32, (one, one), (1 << 22) + 1, 8, 4, 6144, 8)]:
        st, sz = _size(fn, *args, n)
        assert st == 0
        assert sz == radix_aux(n, kb, vb, nv, passes), (fn, n)
    st, sz = _size(lib.mgpu_scan_exclusive_i32, one, 1 << 20, 0)
    assert st == 0 and sz == r256(8 * -(-(1 << 20) // (256 * 80)))  # 8 B per 20480-element tile


def test_invalid_arguments():
    lib = mgpu_hip.load()
    one = ctypes.c_void_p(1)
    st, _ = _size(lib.mgpu_mergesort_keys_u32, one, -1, 0)
    assert st == mgpu_hip.MGPU_INVALID_ARG
    # Sizing does not dereference data pointers (the reference sizes with
    # NULL data, vk/radix.hxx:299); the real call rejects them.
    st, _ = _size(lib.mgpu_mergesort_keys_u32, None, 10000, 0)
    assert st == mgpu_hip.MGPU_OK
    cap = ctypes.c_size_t(1 << 20)
    assert lib.mgpu_mergesort_keys_u32(one, ctypes.byref(cap), None, None, 10000, 0) == \
        mgpu_hip.MGPU_INVALID_ARG
    st, _ = _size(lib.mgpu_mergesort_keys_u32, one, 10, 7)
    assert st == mgpu_hip.MGPU_INVALID_ARG
    st, _ = _size(lib.mgpu_radix_sort_keys_u32, one, 1 << 32)
    assert st == mgpu_hip.MGPU_UNSUPPORTED
    st, _ = _size(lib.mgpu_scan_i32, one, one, 10, 5, 0)
    assert st == mgpu_hip.MGPU_INVALID_ARG
    # aux_size pointer is mandatory.
    assert lib.mgpu_scan_exclusive_i32(None, None, None, one, 10, 0) == mgpu_hip.MGPU_INVALID_ARG


def test_count_zero_is_noop():
    lib = mgpu_hip.load()
    for fn, args in [(lib.mgpu_mergesort_keys_f32, (None, 0, 0)),
                     (lib.mgpu_radix_sort_keys_u64, (None, 0)),
                     (lib.mgpu_scan_exclusive_f32, (None, 0, 0.0)),
                     (lib.mgpu_merge_keys_i32, (None, 0, None, 0, None, 0))]:
        st, sz = _size(fn, *args)
        assert st == 0 and sz == 0


def test_aux_too_small_is_reported_before_any_launch():
    lib = mgpu_hip.load()
    one = ctypes.c_void_p(1)
    cap = ctypes.c_size_t(16)
    st = lib.mgpu_mergesort_keys_u32(one, ctypes.byref(cap), None, one, 1 << 20, 0)
    assert st == mgpu_hip.MGPU_AUX_TOO_SMALL
    cap = ctypes.c_size_t(16)
    st = lib.mgpu_radix_sort_keys_u32(one, ctypes.byref(cap), None, one, 1 << 20)
    assert st == mgpu_hip.MGPU_AUX_TOO_SMALL


def test_misaligned_aux_is_rejected_before_any_launch():
    """aux is carved at 256-byte offsets and holds atomically updated status
    words, so an aux that is not 256-byte aligned is an argument error (after
    the capacity check).  Data pointers are never dereferenced here."""
    lib = mgpu_hip.load()
    one = ctypes.c_void_p(1)
    seg = ctypes.c_void_p(256)
    calls = [
        (lib.mgpu_mergesort_keys_u32, (one, 1 << 20, 0)),
        (lib.mgpu_merge_keys_u32, (one, 1 << 20, one, 1 << 20, one, 0)),
        (lib.mgpu_radix_sort_kv_u32_u32, (one, one, 1 << 20)),
        (lib.mgpu_scan_i32, (one, one, 1 << 20, 0, 0)),
        (lib.mgpu_reduce_i32, (one, 1 << 20, one, 0)),
        (lib.mgpu_scan_u64, (one, one, 1 << 20, 0, 0)),
        (lib.mgpu_segmented_scan_i32, (one, one, 1 << 20, seg, 100, 0, 0)),
        (lib.mgpu_segmented_sort_keys_u32, (one, 1 << 20, seg, 100, 0, None)),
    ]
    for fn, args in calls:
        st, need = _size(fn, *args)
        assert st == mgpu_hip.MGPU_OK and need > 0, fn
        for base in (0x100000 + 4, 0x100000 + 128):
            cap = ctypes.c_size_t(need)
            assert fn(ctypes.c_void_p(base), ctypes.byref(cap), None, *args) == \
                mgpu_hip.MGPU_INVALID_ARG, (fn, base)
            cap = ctypes.c_size_t(16)
            assert fn(ctypes.c_void_p(base), ctypes.byref(cap), None, *args) == \
                mgpu_hip.MGPU_AUX_TOO_SMALL, (fn, base)


def test_segsort_tuning_tables_are_checked_before_any_launch():
    """A bucket table is checked by the sizing call: a blocksort shape whose
    tile exceeds the 160 KB of LDS for the key and value types (1024 x 17 with
    8-byte keys and values), more buckets than the table holds, or capacities
    that do not ascend are argument errors; the default table of each type is
    accepted.  No pointer is dereferenced."""
    lib = mgpu_hip.load()
    seg = ctypes.c_void_p(256)
    shapes = [(nt, vt) for nt, vt in zip(*_shape_arrays(lib))]
    big = shapes.index((1024, 17))
    caps = [128, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 16384]
    sh = [8, 0, 9, 1, 10, 2, 11, 4, 13, 5, 14, 7, big]
    for i, c in enumerate(caps):
        assert shapes[sh[i]][0] * shapes[sh[i]][1] >= c

    def table(cs, ss, nb=None):
        t = mgpu_hip._SegsortTuning()
        t.num_buckets = len(cs) if nb is None else nb
        for b, (c, x) in enumerate(zip(cs, ss)):
            t.cap[b] = c
            t.shape[b] = x
        return ctypes.byref(t)

    def size(fn, tuning, kv):
        args = (ctypes.c_void_p(1),) * (2 if kv else 1) + (1 << 20, seg, 100, 0, tuning)
        return _size(fn, *args)[0]

    full = table(caps, sh)
    assert size(lib.mgpu_segmented_sort_kv_u32_u32, full, True) == mgpu_hip.MGPU_OK
    assert size(lib.mgpu_segmented_sort_keys_u64, full, False) == mgpu_hip.MGPU_OK
    assert size(lib.mgpu_segmented_sort_kv_u64_u32, full, True) == mgpu_hip.MGPU_INVALID_ARG
    assert size(lib.mgpu_segmented_sort_kv_f64_u32, full, True) == mgpu_hip.MGPU_INVALID_ARG
    assert size(lib.mgpu_segmented_sort_kv_u64_u32, table(caps[:12], sh[:12]), True) == \
        mgpu_hip.MGPU_OK
    assert size(lib.mgpu_segmented_sort_kv_u32_u32, table(caps, sh, nb=14), True) == \
        mgpu_hip.MGPU_INVALID_ARG
    assert size(lib.mgpu_segmented_sort_kv_u32_u32, table(caps[::-1], sh[::-1]), True) == \
        mgpu_hip.MGPU_INVALID_ARG
    for fn, kv in [(lib.mgpu_segmented_sort_kv_u32_u32, True),
                   (lib.mgpu_segmented_sort_kv_u64_u32, True),
                   (lib.mgpu_segmented_sort_keys_f64, False)]:
        assert size(fn, None, kv) == mgpu_hip.MGPU_OK


def _shape_arrays(lib):
    nt = (ctypes.c_int * 64)()
    vt = (ctypes.c_int * 64)()
    n = lib.mgpu_segsort_shapes(nt, vt, 64)
    return list(nt[:n]), list(vt[:n])


def test_python_surface_rejects_host_tensors():
    import torch
    with pytest.raises(ValueError):
        mgpu_hip.mergesort_keys(torch.zeros(10, dtype=torch.int32))


def test_python_merge_rejects_mismatched_shapes():
    """The C ABI takes raw pointers and the counts of a and b, so the Python
    surface checks that c and every value array match those counts (and that
    the key dtypes agree) before anything reaches the device."""
    import torch
    k = lambda n, dt=torch.uint32: torch.zeros(n, dtype=dt)
    with pytest.raises(ValueError, match="b_vals"):
        mgpu_hip.merge(k(5), k(5), k(3), k(4), k(8), k(8), "less")
    with pytest.raises(ValueError, match="c_vals"):
        mgpu_hip.merge(k(5), k(5), k(3), k(3), k(8), k(7), "less")
    with pytest.raises(ValueError, match="c must hold"):
        mgpu_hip.merge_keys(k(5), k(3), k(9), "less")
    with pytest.raises(ValueError, match="one dtype"):
        mgpu_hip.merge_keys(k(5), k(3, torch.uint64), k(8), "less")


def test_python_scan_and_reduce_reject_mismatched_element_sizes():
    """Scans write sizeof(type) bytes per element: an output narrower than the
    input, an elem_type that does not match the tensor, or an empty reduce
    output would be overrun, so the Python surface refuses them."""
    import torch
    with pytest.raises(ValueError, match="output element size"):
        mgpu_hip.scan_into(torch.zeros(8, dtype=torch.uint64), torch.zeros(8, dtype=torch.int32))
    with pytest.raises(ValueError, match="input element size"):
        mgpu_hip.scan(torch.zeros(8, dtype=torch.int32), elem_type="u64")
    with pytest.raises(ValueError, match="elem_type"):
        mgpu_hip.scan(torch.zeros(8, dtype=torch.int32), elem_type="i16")
    with pytest.raises(ValueError, match="one element"):
        mgpu_hip.reduce(torch.zeros(8, dtype=torch.int32), torch.zeros(0, dtype=torch.int32))
    with pytest.raises(ValueError, match="output element size"):
        mgpu_hip.reduce(torch.zeros(8, dtype=torch.uint64), torch.zeros(1, dtype=torch.int32))
    with pytest.raises(ValueError, match="output element size"):
        mgpu_hip.segmented_scan(torch.zeros(8, dtype=torch.float32),
                                torch.zeros(8, dtype=torch.float64),
                                torch.zeros(1, dtype=torch.int32))
