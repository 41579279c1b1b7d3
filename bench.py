#!/usr/bin/env python3
"""Benchmark of the sort / merge / scan hot path on one MI355X.

Headline step: one stable key-value LSD radix sort (8-bit digits, 4 digit
passes of count + scan + scatter) of 2^28 uint32 keys drawn uniform from SplitMix64
seed 0, with uint32 values set to the original index.  Inputs are generated on
the device and restored by a device-to-device copy before every step; each
step's sort is bracketed by HIP events on the launch stream, so the timed
region starts with the inputs resident in HBM and excludes the restore.

The JSON line also carries:
  roofline      the dominant kernel (the scatter of one digit pass): algorithmic
                bytes per launch (16 B/key: read key + value, write key +
                value) / its average duration from the library's per-launch
                HIP events, against the 8 TB/s HBM peak; `traffic` is read
                from profiles/traffic.json when a rocprofv3 PMC measurement
                of the same kernel has been recorded there; `copy_ceiling`
                is a plain 2 GiB device-to-device copy by the library's copy
                kernel (and by the HIP runtime) timed the same way: context
                for the memory-bound kernels, not the denominator;
  cpu_baseline  std::stable_sort of (key, index) pairs from the CPU oracle,
                one core, on a bounded sample of the same distribution;
  cpu_baseline_parallel  the same pairs sorted by up to 16 std::threads
                (chunk std::stable_sort + merge-path-split std::merge rounds),
                labelled separately;
  segsort       65536 lognormal segments (~2^25.6 keys): batched segmented
                sort (eager and graph replay) vs a loop of plain sorts;
  secondary     the other primitives at their workload sizes (mergesort kv
                2^28, merge 2 x 2^27, scans / reduce 2^29, low-entropy radix,
                segmented scan), and key-only radix at the sizes of the
                reference's published radix chart (10^6 x {1 ... 50} keys);
  vendor_rocprim  AMD's rocPRIM (compiled from ROCm's source headers into
                vendor_rocprim.so) on the same inputs in the same process:
                radix sort and merge sort of the 2^28 pairs, radix sort of
                2^27 u64 pairs and of 10^6 / 5 x 10^7 u32 keys, merge, scans
                (i32, f32, u64), reduce and the segmented sort layout, with
                each pair of outputs compared where the result is unique;
  serving       the request-serving engine at saturation and three loads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-secondary]
                       [--no-cpu] [--no-serving] [--no-segsort] [--no-segsort-loop] [--no-chart]
                       [--no-vendor] [--cpu-log2 L]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mgpu-shaders_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(torch, steps, warmup, restore, run):
    """W untimed steps, then K steps each timed by events around `run` only."""
    for _ in range(warmup):
        restore()
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        restore()
        e0.record()
        run()
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    return sum(ms) / len(ms), wall * 1000.0 / steps


def kernel_times(mgpu_hip, torch, steps, restore, run):
    """{kernel: avg ms per launch} from the library's per-launch events."""
    restore()
    run()
    torch.cuda.synchronize()
    mgpu_hip.timing_enable(True)
    for _ in range(steps):
        restore()
        run()
    torch.cuda.synchronize()
    t = mgpu_hip.timing_collect()
    mgpu_hip.timing_enable(False)
    return {k: (n, tot / n) for k, (n, tot) in t.items()}


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def cpu_baseline(log2n, reps=3):
    """std::stable_sort of (key, index) pairs on one core (the oracle): the
    median of `reps` runs over the same 2^log2n pairs (input generation and
    the copy the oracle sorts are outside the timing)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import oracle
    n = 1 << log2n
    keys = oracle.generate(n, "u32", 0)
    vals = np.arange(n, dtype=np.uint32)
    ts = [oracle.stable_sort_timed(keys, vals, "u32") for _ in range(reps)]
    dt = sorted(ts)[len(ts) // 2]
    return {"value": n / dt, "unit": "keys/s", "cores": 1, "kind": "port",
            "sample": "std::stable_sort of 2^%d (u32 key, u32 index) pairs, uniform keys "
                      "(seed 0), median of %d runs (%s s) on one core, nproc=%d"
                      % (log2n, reps, ", ".join("%.2f" % t for t in ts), os.cpu_count() or 0),
            "seconds": ts, "cpu_model": cpu_model()}


def cpu_baseline_parallel(log2n, reps=3, threads=None):
    """A separately labelled multi-threaded CPU baseline (the oracle): the
    same 2^log2n (key, index) pairs stable-sorted by `threads` std::threads,
    one std::stable_sort per chunk, then merge rounds each split across the
    threads by merge path; median of `reps` runs."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import oracle
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    n = 1 << log2n
    keys = oracle.generate(n, "u32", 0)
    vals = np.arange(n, dtype=np.uint32)
    ts = [oracle.stable_sort_threads(keys, vals, "u32", False, threads)[1] for _ in range(reps)]
    dt = sorted(ts)[len(ts) // 2]
    return {"value": n / dt, "unit": "keys/s", "cores": threads, "kind": "port",
            "sample": "2^%d (u32 key, u32 index) pairs, uniform keys (seed 0), %d std::threads: "
                      "std::stable_sort per chunk, then std::merge rounds split by merge path; "
                      "median of %d runs (%s s), nproc=%d"
                      % (log2n, threads, reps, ", ".join("%.2f" % t for t in ts),
                         os.cpu_count() or 0),
            "seconds": ts}


def copy_ceiling(torch, mgpu_hip, steps=10):
    """Plain device-to-device copies of 2 GiB timed with events: the library's
    HBM copy kernel (16 KB tiles, non-temporal loads; the practical read +
    write ceiling the memory-bound kernels are compared with) and the HIP
    runtime's copy (torch copy_) beside it.  Fractions stay against the
    8 TB/s peak."""
    n = 1 << 29
    src = torch.empty(n, dtype=torch.uint32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return sorted(ms)[len(ms) // 2]

    moved = 2 * 4 * n
    t = timed(lambda: mgpu_hip.copy(dst, src))
    ok = bool(torch.equal(dst, src))
    t_rt = timed(lambda: dst.copy_(src))
    del src, dst
    return {"bytes_moved": moved, "ms": t, "GBps": moved / (t / 1e3) / 1e9,
            "frac_of_peak": moved / (t / 1e3) / 1e9 / HBM_PEAK_GBS, "copy_correct": ok,
            "runtime_copy_ms": t_rt, "runtime_copy_GBps": moved / (t_rt / 1e3) / 1e9,
            "what": "copy of 2^29 uint32 (2 GiB read + 2 GiB written) by mgpu_copy, and by the "
                    "HIP runtime (torch copy_) for comparison; medians of %d" % steps}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines_other(log2n=27):
    """The reference's other CPU comparators on one core (the oracle):
    std::merge, std::exclusive_scan, std::inclusive_scan (float32) and
    std::reduce over 2^log2n elements, median of 3."""
    sys.path.insert(0, ROOT)
    import oracle
    n = 1 << log2n

    def med(fn):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[1]

    import numpy as np
    a = np.sort(oracle.generate(n // 2, "u32", 1))
    b = np.sort(oracle.generate(n // 2, "u32", 2))
    xi = oracle.generate(n, "i32_0_16", 0)
    xf = oracle.generate(n, "f32_01", 0)
    out = {}
    dt = med(lambda: oracle.merge(a, b))
    out["std_merge_u32"] = {"GBps": 8 * n / dt / 1e9, "keys_per_s": n / dt, "s": dt,
                            "sample": "2 x 2^%d sorted u32 keys" % (log2n - 1)}
    dt = med(lambda: oracle.exclusive_scan_u32(xi))
    out["std_exclusive_scan_i32"] = {"GBps": 8 * n / dt / 1e9, "s": dt,
                                     "sample": "2^%d int32 in [0, 16)" % log2n}
    dt = med(lambda: oracle.inclusive_scan_f32(xf))
    out["std_inclusive_scan_f32"] = {"GBps": 8 * n / dt / 1e9, "s": dt,
                                     "sample": "2^%d float32 in [0, 1)" % log2n}
    dt = med(lambda: oracle.reduce_u32(xi))
    out["std_reduce_i32"] = {"GBps": 4 * n / dt / 1e9, "s": dt,
                             "sample": "2^%d int32 in [0, 16)" % log2n}
    out["cores"] = 1
    out["kind"] = "port"
    return out


def secondary(torch, mgpu_hip, steps, warmup, chart=True):
    out = {}
    dev = "cuda"
    n = 1 << 28

    def gen(count, kind, seed=0, offset=0, dtype=torch.uint32):
        t = torch.empty(count, dtype=dtype, device=dev)
        mgpu_hip.generate(t, kind, seed, offset)
        return t

    # Low-entropy radix and mergesort, 2^28 key-value.
    for name, kind, algo in [("radix_kv_u32_2p28_lowentropy", "lowentropy", "radix"),
                             ("mergesort_kv_u32_2p28_uniform", "u32", "mergesort")]:
        ks, vs = gen(n, kind), gen(n, "iota")
        k, v = torch.empty_like(ks), torch.empty_like(vs)

        def restore():
            k.copy_(ks)
            v.copy_(vs)
        if algo == "radix":
            aux = mgpu_hip.radix_sort(k, v)
            run = lambda: mgpu_hip.radix_sort(k, v, aux=aux)
            kname = "mgpu_radix_scatter"
        else:
            aux = mgpu_hip.mergesort_kv(k, v)
            run = lambda: mgpu_hip.mergesort_kv(k, v, aux=aux)
            kname = "mgpu_mergesort_pass"
        ms, _ = timed_steps(torch, steps, warmup, restore, run)
        kt = kernel_times(mgpu_hip, torch, 2, restore, run)
        per = kt.get(kname, (0, float("nan")))[1]
        out[name] = {"keys_per_s": n / (ms / 1e3), "ms": ms,
                     "pass_kernel": kname, "pass_ms": per,
                     "pass_GBps": 16 * n / (per / 1e3) / 1e9,
                     "pass_frac": 16 * n / (per / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "kernels_ms": {k2: round(v2[1], 4) for k2, v2 in kt.items()}}
        log(name, out[name]["keys_per_s"] / 1e9, "Gkeys/s, pass", per, "ms")
        del ks, vs, k, v, aux

    # 2^20 uint32 keys, key-only mergesort (a latency-sized case: fits in L2 / MALL).
    c1 = 1 << 20
    ks = gen(c1, "u32", 3)
    k = torch.empty_like(ks)
    aux = mgpu_hip.mergesort_keys(k)
    ms, _ = timed_steps(torch, steps, warmup, lambda: k.copy_(ks),
                        lambda: mgpu_hip.mergesort_keys(k, aux=aux))
    out["mergesort_keys_u32_2p20"] = {"keys_per_s": c1 / (ms / 1e3), "ms": ms}
    log("mergesort_keys_u32_2p20", out["mergesort_keys_u32_2p20"])
    del ks, k, aux

    # Key-only uint32 radix sort at the sizes of the reference's only published
    # benchmark (examples/vk_radix/vk_radix.cxx:40-71: 10^6 x s keys), next to
    # the chart's readings for it (examples/vk_radix/timings.jpg, RTX 2060,
    # Mkeys/s read off the chart as a range).  Keys are restored to uniform
    # random before every sort; each sort is timed alone with events.
    chart = {1: (2530, 2630), 2: (3570, 3670), 4: (3960, 4060), 6: (3960, 4060),
             8: (3960, 4060), 50: (3850, 3950)}
    rows = []
    for s_m in (1, 2, 4, 6, 8, 10, 12, 14, 16, 20, 25, 30, 35, 40, 45, 50) if chart else ():
        cnt = 1000000 * s_m
        ks = gen(cnt, "u32", 4)
        k = torch.empty_like(ks)
        aux = mgpu_hip.radix_sort(k)
        ms, _ = timed_steps(torch, max(steps, 20), warmup, lambda: k.copy_(ks),
                            lambda: mgpu_hip.radix_sort(k, aux=aux))
        lo, hi = chart.get(s_m, (4110, 4200))
        rows.append({"keys": cnt, "ms": ms, "mkeys_per_s": cnt / (ms / 1e3) / 1e6,
                     "chart_mkeys_per_s": [lo, hi],
                     "ratio_vs_chart_high": cnt / (ms / 1e3) / 1e6 / hi})
        del ks, k, aux
    if rows:
        out["radix_keys_u32_reference_chart_sizes"] = {
            "rows": rows,
            "algorithm": "count + scan + scatter per digit; tile shape by key count",
            "min_ratio_vs_chart_high": min(r["ratio_vs_chart_high"] for r in rows)}
        log("radix key-only at chart sizes",
            [(r["keys"] // 1000000, round(r["mkeys_per_s"]), round(r["ratio_vs_chart_high"], 1))
             for r in rows])

    # Merge of two sorted 2^27 uint32 arrays, key-only.
    m = 1 << 27
    a, b = gen(m, "u32", 1), gen(m, "u32", 2)
    mgpu_hip.radix_sort(a)
    mgpu_hip.radix_sort(b)
    c = torch.empty(2 * m, dtype=torch.uint32, device=dev)
    aux = mgpu_hip.merge_keys(a, b, c, "less")
    run = lambda: mgpu_hip.merge_keys(a, b, c, "less", aux=aux)
    ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    kt = kernel_times(mgpu_hip, torch, 2, lambda: None, run)
    per = kt.get("mgpu_merge", (0, float("nan")))[1]
    out["merge_keys_u32_2x2p27"] = {"GBps": 8 * 2 * m / (ms / 1e3) / 1e9, "ms": ms,
                                    "merge_kernel_ms": per,
                                    "frac": 8 * 2 * m / (per / 1e3) / 1e9 / HBM_PEAK_GBS,
                                    "kernels_ms": {k2: round(v2[1], 4) for k2, v2 in kt.items()}}
    log("merge", out["merge_keys_u32_2x2p27"])
    del a, b, c, aux

    # Scans and reduces of 2^29 elements.
    s = 1 << 29
    xi = gen(s, "i32_0_16", 0, dtype=torch.int32)
    yi = torch.empty_like(xi)
    xf = gen(s, "f32_01", 0, dtype=torch.float32)
    yf = torch.empty_like(xf)
    r = torch.zeros(1, dtype=torch.int32, device=dev)
    rf = torch.zeros(1, dtype=torch.float32, device=dev)
    cases = [
        ("scan_exclusive_i32_2p29", lambda aux: mgpu_hip.scan_into(xi, yi, aux=aux), 8 * s,
         "mgpu_scan_lookback"),
        ("scan_inclusive_f32_2p29",
         lambda aux: mgpu_hip.scan_into(xf, yf, inclusive=True, aux=aux), 8 * s,
         "mgpu_scan_lookback"),
        ("reduce_i32_2p29", lambda aux: mgpu_hip.reduce(xi, r, aux=aux), 4 * s,
         "mgpu_reduce_partial"),
        ("reduce_f32_2p29", lambda aux: mgpu_hip.reduce(xf, rf, aux=aux), 4 * s,
         "mgpu_reduce_partial"),
        ("segmented_scan_exclusive_i32_2p29_65536_segments",
         lambda aux: mgpu_hip.segmented_scan(xi, yi, segs, aux=aux), 8 * s, "mgpu_segscan"),
        ("scan_exclusive_u64_2p27", lambda aux: mgpu_hip.scan_into(x64, y64, aux=aux),
         16 * (s // 4), "mgpu_scan64_lookback"),
    ]
    # 2^27 uint64 over the first half of the int32 input's bytes (a view).
    x64, y64 = xi.view(torch.uint64)[: s // 4], yi.view(torch.uint64)[: s // 4]
    from mgpu_hip import workloads
    lens = workloads.lognormal_lengths(65536, 2048.0, 1.5, 1, 1 << 22, seed=0)
    starts = workloads.segment_starts(lens)
    segs = torch.from_numpy(starts[starts < s]).to(dev)
    for name, fn, nbytes, kname in cases:
        aux = fn(None)
        run = lambda fn=fn, aux=aux: fn(aux)
        ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
        kt = kernel_times(mgpu_hip, torch, 2, lambda: None, run)
        per = kt.get(kname, (0, float("nan")))[1]
        out[name] = {"GBps": nbytes / (ms / 1e3) / 1e9, "ms": ms, "kernel_ms": per,
                     "frac": nbytes / (per / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "kernels_ms": {k2: round(v2[1], 4) for k2, v2 in kt.items()}}
        log(name, out[name])
    return out


def segsort_loop_baseline(torch, mgpu_hip, k, v, restore, starts, lens, batched_k, batched_v):
    """One plain mergesort call per segment (segments of length <= 1 are
    skipped), through the C ABI with a pre-sized aux buffer: as Python calls,
    and recorded once into a graph and replayed (the GPU-side cost of one
    launch sequence per segment, without the host dispatch of each call).
    Both results are compared with the batched sort's on the device."""
    import ctypes
    lib = mgpu_hip.load()
    fn = lib.mgpu_mergesort_kv_u32_u32
    big = int(lens.max())
    aux_loop = mgpu_hip.mergesort_kv(k[:big], v[:big])
    cap = ctypes.c_size_t(aux_loop.numel())
    auxp = ctypes.c_void_p(aux_loop.data_ptr())
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    kb, vb = k.data_ptr(), v.data_ptr()
    work = [(ctypes.c_void_p(kb + 4 * int(s)), ctypes.c_void_p(vb + 4 * int(s)), int(l))
            for s, l in zip(starts, lens) if l > 1]
    log("segsort loop baseline: %d plain sorts" % len(work))
    restore()
    for kp, vp, l in work[:200]:
        fn(auxp, ctypes.byref(cap), sp, kp, vp, l, 0)
    restore()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for kp, vp, l in work:
        st = fn(auxp, ctypes.byref(cap), sp, kp, vp, l, 0)
        if st:
            raise RuntimeError("plain mergesort failed: %d" % st)
    torch.cuda.synchronize()
    ms_loop = (time.perf_counter() - t0) * 1e3
    same = bool(torch.equal(k.view(torch.int32), batched_k.view(torch.int32)) and
                torch.equal(v.view(torch.int32), batched_v.view(torch.int32)))
    log("segsort loop baseline: %.1f ms as calls; recording the graph" % ms_loop)

    lstream = torch.cuda.Stream()
    torch.cuda.synchronize()
    lsp = ctypes.c_void_p(lstream.cuda_stream)
    with torch.cuda.stream(lstream):
        lg = mgpu_hip.Graph(lstream)
        with lg:
            for kp, vp, l in work:
                st = fn(auxp, ctypes.byref(cap), lsp, kp, vp, l, 0)
                if st:
                    raise RuntimeError("plain mergesort failed: %d" % st)
        ms_loop_graph, _ = timed_steps(torch, 3, 1, restore, lg.launch)
    torch.cuda.synchronize()
    same_graph = bool(torch.equal(k.view(torch.int32), batched_k.view(torch.int32)) and
                      torch.equal(v.view(torch.int32), batched_v.view(torch.int32)))
    del lg
    return {"ms_loop_of_plain_sorts": ms_loop, "plain_sort_calls": len(work),
            "ms_loop_of_plain_sorts_graph": ms_loop_graph,
            "outputs_identical": same and same_graph}


def bench_segsort(torch, mgpu_hip, steps, warmup, loop_baseline=True):
    """65536 lognormal segments (median 256, sigma 1.5, clamped to [1, 2^20]),
    uint32 keys + uint32 values: batched segmented sort (eager and replayed
    as one graph) against a loop of plain mergesort calls, one per segment,
    on the same GPU; outputs compared on the device.  loop_baseline=False
    leaves the loop out (its 65k calls crawl under a counter profiler)."""
    import numpy as np
    from mgpu_hip import workloads
    lens = workloads.lognormal_lengths(65536, 256.0, 1.5, 1, 1 << 20, seed=0)
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    dev = "cuda"
    segs = torch.from_numpy(starts).to(dev)
    ks = torch.empty(n, dtype=torch.uint32, device=dev)
    mgpu_hip.generate(ks, "u32", 0)
    vs = torch.arange(n, dtype=torch.int32, device=dev).view(torch.uint32)
    k = torch.empty_like(ks)
    v = torch.empty_like(vs)

    def restore():
        k.copy_(ks)
        v.copy_(vs)

    tuning = mgpu_hip.segmented_sort_autotune("u32", True, "less")
    restore()
    aux = mgpu_hip.segmented_sort(k, segs, v)
    run = lambda: mgpu_hip.segmented_sort(k, segs, v, aux=aux)
    ms_eager, _ = timed_steps(torch, steps, warmup, restore, run)
    kt = kernel_times(mgpu_hip, torch, 2, restore, run)

    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        g = mgpu_hip.Graph(stream)
        with g:
            mgpu_hip.segmented_sort(k, segs, v, aux=aux, stream=stream)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        ms_graph, _ = timed_steps(torch, steps, warmup, restore, g.launch)
    torch.cuda.synchronize()
    batched_k, batched_v = k.clone(), v.clone()
    log("segsort batched: eager %.4f ms, graph %.4f ms" % (ms_eager, ms_graph))
    if loop_baseline:
        loop = segsort_loop_baseline(torch, mgpu_hip, k, v, restore, starts, lens,
                                     batched_k, batched_v)
    else:
        loop = {"ms_loop_of_plain_sorts": None, "plain_sort_calls": None,
                "ms_loop_of_plain_sorts_graph": None, "outputs_identical": None}

    # Algorithmic bytes: 16 B per key-value pair for every pass actually run:
    # one read and one write of the segments the blocksort buckets hold, four
    # digit passes over the segments longer than the largest bucket; segments
    # of one key are not touched.
    cap_max = max(tuning["cap"])
    small = int(lens[(lens > 1) & (lens <= cap_max)].sum())
    large = int(lens[lens > cap_max].sum())
    alg_bytes = 16 * small + 16 * 4 * large
    shapes = mgpu_hip.segsort_shapes()
    out = {"segments": 65536, "keys": n, "log2_keys": float(np.log2(n)),
           "ms_batched_eager": ms_eager, "ms_batched_graph": ms_graph,
           "keys_per_s_graph": n / (ms_graph / 1e3),
           "roofline": {"bound": "hbm", "algorithmic_bytes_per_call": alg_bytes,
                        "keys_in_buckets": small, "keys_in_large_segments": large,
                        "achieved": alg_bytes / (ms_graph / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s",
                        "frac": alg_bytes / (ms_graph / 1e3) / 1e9 / HBM_PEAK_GBS},
           **loop,
           "speedup_vs_loop": (loop["ms_loop_of_plain_sorts"] / ms_graph
                               if loop_baseline else None),
           "speedup_vs_loop_graph": (loop["ms_loop_of_plain_sorts_graph"] / ms_graph
                                     if loop_baseline else None),
           "tuning": {"cap": tuning["cap"], "shape": [list(shapes[s]) for s in tuning["shape"]]},
           "kernels_ms": {k2: round(v2[1], 4) for k2, v2 in kt.items()}}
    log("segsort", {k2: out[k2] for k2 in ("ms_batched_eager", "ms_batched_graph",
                                           "ms_loop_of_plain_sorts",
                                           "ms_loop_of_plain_sorts_graph", "speedup_vs_loop",
                                           "speedup_vs_loop_graph", "outputs_identical")},
        "frac", out["roofline"]["frac"])
    return out


def bench_vendor(torch, mgpu_hip, steps, warmup):
    """AMD's rocPRIM (vendor_rocprim.so, compiled from ROCm's rocPRIM source
    headers) against mgpu-hip on the same device inputs in the same process,
    each call timed alone with events on the current stream.  Stable sorts,
    the merge, the integer scan and the reduce have unique results, so the
    two outputs are also compared bit for bit."""
    import ctypes
    import numpy as np
    from mgpu_hip import workloads
    path = os.path.join(ROOT, "vendor_rocprim.so")
    if not os.path.exists(path):
        return {"skipped": "vendor_rocprim.so not built"}
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    i64 = ctypes.c_int64
    sz = ctypes.POINTER(ctypes.c_size_t)
    sig = {"vendor_radix_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, vp,
                                           ctypes.POINTER(ctypes.c_int)],
           "vendor_radix_sort_pairs_u64": [vp, sz, vp, vp, vp, vp, i64, vp,
                                           ctypes.POINTER(ctypes.c_int)],
           "vendor_radix_sort_keys_u32": [vp, sz, vp, vp, i64, vp, ctypes.POINTER(ctypes.c_int)],
           "vendor_exclusive_scan_u64": [vp, sz, vp, vp, i64, ctypes.c_uint64, vp],
           "vendor_merge_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, vp],
           "vendor_merge_u32": [vp, sz, vp, i64, vp, i64, vp, vp],
           "vendor_exclusive_scan_i32": [vp, sz, vp, vp, i64, ctypes.c_int32, vp],
           "vendor_inclusive_scan_f32": [vp, sz, vp, vp, i64, vp],
           "vendor_reduce_i32": [vp, sz, vp, i64, vp, vp],
           "vendor_segmented_radix_sort_pairs_u32": [vp, sz, vp, vp, vp, vp, i64, i64, vp, vp,
                                                     vp]}
    for name, args in sig.items():
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = ctypes.c_int
    lib.vendor_rocprim_version.restype = ctypes.c_int
    dev = "cuda"
    p = lambda t: vp(t.data_ptr())
    stream = lambda: vp(torch.cuda.current_stream().cuda_stream)

    def two_phase(fn, args, tail=()):
        """Size rocPRIM's temporary storage, allocate it, return a runner."""
        need = ctypes.c_size_t(0)
        e = fn(None, ctypes.byref(need), *args, stream(), *tail)
        if e:
            raise RuntimeError("rocPRIM sizing failed: %d" % e)
        temp = torch.empty(max(need.value, 1), dtype=torch.uint8, device=dev)
        cap = ctypes.c_size_t(temp.numel())

        def run():
            cap.value = temp.numel()
            e2 = fn(p(temp), ctypes.byref(cap), *args, stream(), *tail)
            if e2:
                raise RuntimeError("rocPRIM call failed: %d" % e2)
        return run, temp

    def same(x, y):
        return bool(torch.equal(x.view(torch.int32), y.view(torch.int32)))

    out = {"library": "rocPRIM %d.%d.%d" % (lib.vendor_rocprim_version() // 100000,
                                            lib.vendor_rocprim_version() // 100 % 1000,
                                            lib.vendor_rocprim_version() % 100),
           "method": "same inputs, same process; each call timed alone (events on the "
                     "stream); 'speedup' = rocPRIM ms / mgpu-hip ms"}

    def row(name, ours_ms, vendor_ms, identical, units, per):
        out[name] = {"mgpu_hip_ms": ours_ms, "rocprim_ms": vendor_ms,
                     "speedup": vendor_ms / ours_ms, "outputs_identical": identical,
                     "mgpu_hip_" + per: units / (ours_ms / 1e3),
                     "rocprim_" + per: units / (vendor_ms / 1e3)}
        log("vs rocPRIM", name, {k: (round(v, 4) if isinstance(v, float) else v)
                                 for k, v in out[name].items()})

    # Radix sort of 2^28 (u32 key, u32 index) pairs; rocPRIM's double-buffer
    # form over the same footprint (keys, values and one alternate pair).
    n = 1 << 28
    ks = torch.empty(n, dtype=torch.uint32, device=dev)
    mgpu_hip.generate(ks, "u32", 0)
    vs = torch.empty_like(ks)
    mgpu_hip.generate(vs, "iota")
    k, v = torch.empty_like(ks), torch.empty_like(vs)
    k2, v2 = torch.empty_like(ks), torch.empty_like(vs)

    def restore():
        k.copy_(ks)
        v.copy_(vs)
    restore()
    aux = mgpu_hip.radix_sort(k, v)
    ours_ms, _ = timed_steps(torch, steps, warmup, restore,
                             lambda: mgpu_hip.radix_sort(k, v, aux=aux))
    ok, ov = k.clone(), v.clone()
    del aux
    alt = ctypes.c_int(0)
    run, temp = two_phase(lib.vendor_radix_sort_pairs_u32, (p(k), p(k2), p(v), p(v2), n),
                          (ctypes.byref(alt),))
    vendor_ms, _ = timed_steps(torch, steps, warmup, restore, run)
    torch.cuda.synchronize()
    rk, rv = (k2, v2) if alt.value else (k, v)
    row("radix_sort_kv_u32_2p28", ours_ms, vendor_ms, same(ok, rk) and same(ov, rv), n,
        "keys_per_s")
    del temp, ok, ov

    # Merge sort of the same pairs (rocPRIM out of place into k2, v2).
    restore()
    aux = mgpu_hip.mergesort_kv(k, v)
    ours_ms, _ = timed_steps(torch, steps, warmup, restore,
                             lambda: mgpu_hip.mergesort_kv(k, v, aux=aux))
    del aux
    run, temp = two_phase(lib.vendor_merge_sort_pairs_u32, (p(k), p(k2), p(v), p(v2), n))
    vendor_ms, _ = timed_steps(torch, steps, warmup, restore, run)
    torch.cuda.synchronize()
    restore()
    mgpu_hip.mergesort_kv(k, v)
    torch.cuda.synchronize()
    row("mergesort_kv_u32_2p28", ours_ms, vendor_ms, same(k, k2) and same(v, v2), n,
        "keys_per_s")
    del temp, ks, vs, k, v, k2, v2

    # 64-bit keys: radix sort of 2^27 (u64 key, u32 index) pairs.
    n64 = 1 << 27
    ks64 = torch.empty(n64, dtype=torch.uint64, device=dev)
    mgpu_hip.generate(ks64, "u64", 0)
    vs64 = torch.empty(n64, dtype=torch.uint32, device=dev)
    mgpu_hip.generate(vs64, "iota")
    k64, v64 = torch.empty_like(ks64), torch.empty_like(vs64)
    k64b, v64b = torch.empty_like(ks64), torch.empty_like(vs64)

    def restore64():
        k64.copy_(ks64)
        v64.copy_(vs64)
    restore64()
    aux = mgpu_hip.radix_sort(k64, v64)
    ours_ms, _ = timed_steps(torch, steps, warmup, restore64,
                             lambda: mgpu_hip.radix_sort(k64, v64, aux=aux))
    ok, ov = k64.clone(), v64.clone()
    del aux
    run, temp = two_phase(lib.vendor_radix_sort_pairs_u64,
                          (p(k64), p(k64b), p(v64), p(v64b), n64), (ctypes.byref(alt),))
    vendor_ms, _ = timed_steps(torch, steps, warmup, restore64, run)
    torch.cuda.synchronize()
    rk, rv = (k64b, v64b) if alt.value else (k64, v64)
    row("radix_sort_kv_u64_2p27", ours_ms, vendor_ms, same(ok, rk) and same(ov, rv), n64,
        "keys_per_s")
    del temp, ok, ov, ks64, vs64, k64, v64, k64b, v64b

    # Key-only 32-bit radix sort at the reference chart's smallest and
    # largest sizes (10^6 and 5 x 10^7 keys).
    for cnt in (1000000, 50000000):
        kc = torch.empty(cnt, dtype=torch.uint32, device=dev)
        mgpu_hip.generate(kc, "u32", 4)
        kk, kb = torch.empty_like(kc), torch.empty_like(kc)
        rst = lambda: kk.copy_(kc)
        rst()
        aux = mgpu_hip.radix_sort(kk)
        ours_ms, _ = timed_steps(torch, max(steps, 20), warmup, rst,
                                 lambda: mgpu_hip.radix_sort(kk, aux=aux))
        ok = kk.clone()
        run, temp = two_phase(lib.vendor_radix_sort_keys_u32, (p(kk), p(kb), cnt),
                              (ctypes.byref(alt),))
        vendor_ms, _ = timed_steps(torch, max(steps, 20), warmup, rst, run)
        torch.cuda.synchronize()
        row("radix_sort_keys_u32_%dM" % (cnt // 1000000), ours_ms, vendor_ms,
            same(ok, kb if alt.value else kk), cnt, "keys_per_s")
        del kc, kk, kb, aux, ok, temp

    # Merge of two sorted 2^27-key arrays.
    m = 1 << 27
    a = torch.empty(m, dtype=torch.uint32, device=dev)
    b = torch.empty_like(a)
    mgpu_hip.generate(a, "u32", 1)
    mgpu_hip.generate(b, "u32", 2)
    mgpu_hip.radix_sort(a)
    mgpu_hip.radix_sort(b)
    c = torch.empty(2 * m, dtype=torch.uint32, device=dev)
    c2 = torch.empty_like(c)
    aux = mgpu_hip.merge_keys(a, b, c, "less")
    ours_ms, _ = timed_steps(torch, steps, warmup, lambda: None,
                             lambda: mgpu_hip.merge_keys(a, b, c, "less", aux=aux))
    run, temp = two_phase(lib.vendor_merge_u32, (p(a), m, p(b), m, p(c2)))
    vendor_ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    torch.cuda.synchronize()
    row("merge_keys_u32_2x2p27", ours_ms, vendor_ms, same(c, c2), 8 * 2 * m, "bytes_per_s")
    del a, b, c, c2, aux, temp

    # Scans and reduce of 2^29 elements.
    s = 1 << 29
    xi = torch.empty(s, dtype=torch.int32, device=dev)
    mgpu_hip.generate(xi, "i32_0_16", 0)
    yi, yi2 = torch.empty_like(xi), torch.empty_like(xi)
    aux = mgpu_hip.scan_into(xi, yi)
    ours_ms, _ = timed_steps(torch, steps, warmup, lambda: None,
                             lambda: mgpu_hip.scan_into(xi, yi, aux=aux))
    run, temp = two_phase(lib.vendor_exclusive_scan_i32, (p(xi), p(yi2), s, 0))
    vendor_ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    torch.cuda.synchronize()
    row("scan_exclusive_i32_2p29", ours_ms, vendor_ms, same(yi, yi2), 8 * s, "bytes_per_s")
    del aux, temp
    r, r2 = torch.zeros(1, dtype=torch.int32, device=dev), torch.ones(1, dtype=torch.int32,
                                                                       device=dev)
    aux = mgpu_hip.reduce(xi, r)
    ours_ms, _ = timed_steps(torch, steps, warmup, lambda: None,
                             lambda: mgpu_hip.reduce(xi, r, aux=aux))
    run, temp = two_phase(lib.vendor_reduce_i32, (p(xi), s, p(r2)))
    vendor_ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    torch.cuda.synchronize()
    row("reduce_i32_2p29", ours_ms, vendor_ms, same(r, r2), 4 * s, "bytes_per_s")
    del aux, temp
    # 2^27 uint64 over the int32 input's bytes (a view), exclusive.
    x64, y64, y64b = xi.view(torch.uint64), yi.view(torch.uint64), yi2.view(torch.uint64)
    aux = mgpu_hip.scan_into(x64, y64)
    ours_ms, _ = timed_steps(torch, steps, warmup, lambda: None,
                             lambda: mgpu_hip.scan_into(x64, y64, aux=aux))
    run, temp = two_phase(lib.vendor_exclusive_scan_u64, (p(x64), p(y64b), s // 2, 0))
    vendor_ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    torch.cuda.synchronize()
    row("scan_exclusive_u64_2p28", ours_ms, vendor_ms, same(y64, y64b), 16 * (s // 2),
        "bytes_per_s")
    del aux, temp, x64, y64, y64b, yi, yi2, xi
    xf = torch.empty(s, dtype=torch.float32, device=dev)
    mgpu_hip.generate(xf, "f32_01", 0)
    yf, yf2 = torch.empty_like(xf), torch.empty_like(xf)
    aux = mgpu_hip.scan_into(xf, yf, inclusive=True)
    ours_ms, _ = timed_steps(torch, steps, warmup, lambda: None,
                             lambda: mgpu_hip.scan_into(xf, yf, inclusive=True, aux=aux))
    run, temp = two_phase(lib.vendor_inclusive_scan_f32, (p(xf), p(yf2), s))
    vendor_ms, _ = timed_steps(torch, steps, warmup, lambda: None, run)
    torch.cuda.synchronize()
    # Different summation orders: report each one's error against float64.
    exact = torch.cumsum(xf.double(), 0)
    row("scan_inclusive_f32_2p29", ours_ms, vendor_ms, None, 8 * s, "bytes_per_s")
    out["scan_inclusive_f32_2p29"]["max_rel_err"] = {
        "mgpu_hip": float(((yf.double() - exact).abs() / exact).max()),
        "rocprim": float(((yf2.double() - exact).abs() / exact).max())}
    del aux, temp, xf, yf, yf2, exact

    # Segmented sort on the C4 layout (65536 lognormal segments).
    lens = workloads.lognormal_lengths(65536, 256.0, 1.5, 1, 1 << 20, seed=0)
    starts = workloads.segment_starts(lens)
    n4 = int(lens.sum())
    segs = torch.from_numpy(starts).to(dev)
    ends = torch.from_numpy(np.append(starts[1:], np.int32(n4)).astype(np.int32)).to(dev)
    ks = torch.empty(n4, dtype=torch.uint32, device=dev)
    mgpu_hip.generate(ks, "u32", 0)
    vs = torch.arange(n4, dtype=torch.int32, device=dev).view(torch.uint32)
    k, v = torch.empty_like(ks), torch.empty_like(vs)
    k2, v2 = torch.empty_like(ks), torch.empty_like(vs)
    restore()
    aux = mgpu_hip.segmented_sort(k, segs, v)
    ours_ms, _ = timed_steps(torch, steps, warmup, restore,
                             lambda: mgpu_hip.segmented_sort(k, segs, v, aux=aux))
    run, temp = two_phase(lib.vendor_segmented_radix_sort_pairs_u32,
                          (p(k), p(k2), p(v), p(v2), n4, len(starts), p(segs), p(ends)))
    vendor_ms, _ = timed_steps(torch, steps, warmup, restore, run)
    torch.cuda.synchronize()
    restore()
    mgpu_hip.segmented_sort(k, segs, v, aux=aux)
    torch.cuda.synchronize()
    row("segmented_sort_kv_u32_65536_lognormal", ours_ms, vendor_ms,
        same(k, k2) and same(v, v2), n4, "keys_per_s")
    del ks, vs, k, v, k2, v2, aux, temp
    return out


def bench_serving(torch, mgpu_hip, count=256):
    """Serving engine under an open-loop load: sort / merge / scan requests
    with key types u32 / u64 / f32 and total sizes log-uniform over
    [2^10, 2^26], 8 slots, 64 GiB budget.  Saturation = all requests
    submitted at once; then Poisson arrivals at 0.5x, 0.8x and 1.1x of the
    saturation request rate.  Latency = completion - scheduled arrival.
    Every response is verified after its run (outside the timed window)."""
    from mgpu_hip import serving

    def one(rate, seed):
        reqs = serving.make_requests(count, seed=seed)
        sv = serving.Server(slots=8, budget_bytes=64 << 30)
        try:
            resps = serving.run_load(sv, reqs, rate=rate)
            summ = serving.summarize(resps, reqs)
            worst = 0.0
            for r in resps:
                worst = max(worst, serving.verify(r))
                sv.release(r["id"])
            summ["verified"] = len(resps)
            summ["f32_scan_max_rel_err"] = worst
            summ["peak_budget_gib"] = sv.stats()["peak"] / 2 ** 30
        finally:
            sv.close()
        torch.cuda.synchronize()
        return summ

    one(None, 99)  # warm-up: pool growth, code objects
    sat = one(None, 1)
    log("serving saturation", sat)
    out = {"requests_per_run": count, "slots": 8, "budget_gib": 64,
           "sizes": "total elements log-uniform [2^10, 2^26]", "saturation": sat, "rates": []}
    for i, f in enumerate((0.5, 0.8, 1.1)):
        rate = f * sat["requests_per_s"]
        summ = one(rate, 2 + i)
        summ["load_factor"] = f
        summ["offered_requests_per_s"] = rate
        log("serving", f, {k: summ[k] for k in ("p50_ms", "p99_ms", "keys_per_s")})
        out["rates"].append(summ)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-segsort", action="store_true")
    ap.add_argument("--no-segsort-loop", action="store_true",
                    help="leave out the segmented sort's loop-of-plain-sorts comparison")
    ap.add_argument("--no-chart", action="store_true",
                    help="skip the key-only radix sizes of the reference chart")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip the rocPRIM comparison (vendor_rocprim.so)")
    ap.add_argument("--cpu-log2", type=int, default=27)
    ap.add_argument("--cpu-reps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(--gpus %d requested; run with --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import mgpu_hip

    mgpu_hip.load()
    n = 1 << 28
    dev = "cuda"
    keys_src = torch.empty(n, dtype=torch.uint32, device=dev)
    vals_src = torch.empty(n, dtype=torch.uint32, device=dev)
    mgpu_hip.generate(keys_src, "u32", 0)
    mgpu_hip.generate(vals_src, "iota")
    keys = torch.empty_like(keys_src)
    vals = torch.empty_like(vals_src)

    def restore():
        keys.copy_(keys_src)
        vals.copy_(vals_src)

    restore()
    aux = mgpu_hip.radix_sort(keys, vals)
    run = lambda: mgpu_hip.radix_sort(keys, vals, aux=aux)
    log("radix kv 2^28: warmup", args.warmup, "steps", args.steps)
    mgpu_hip.lookback_fallbacks(reset=True)
    ms, wall_ms = timed_steps(torch, args.steps, args.warmup, restore, run)
    fallbacks = mgpu_hip.lookback_fallbacks(reset=True)
    kt = kernel_times(mgpu_hip, torch, min(args.steps, 5), restore, run)
    log("kernel times", kt)
    launches, pass_ms = kt["mgpu_radix_scatter"]
    pass_bytes = 16 * n
    achieved = pass_bytes / (pass_ms / 1e3) / 1e9
    workload = "radix_sort_kv_u32_2p28_uniform"
    traffic = load_traffic(workload)
    result = {
        "metric": "keys/s, 2^28 uint32 key-value LSD radix sort (8-bit digits)",
        "value": n / (ms / 1e3),
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "wall_ms_per_step_incl_restore": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (SplitMix64 seed 0, generated on device)",
        "config": {"workload": "2^28 uint32 keys uniform over [0,2^32), uint32 values = index, "
                               "stable LSD radix sort, 8-bit digits, 4 passes of "
                               "count + scan + scatter",
                   "n": n, "key": "u32", "value": "u32", "nt": 512, "vt": 20},
        "roofline": {"bound": "hbm", "kernel": "mgpu_radix_scatter",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "algorithmic_bytes_per_launch": pass_bytes,
                     "avg_launch_ms": pass_ms, "launches": launches,
                     "whole_pass_ms": ms / 4,
                     "whole_pass_frac": pass_bytes / (ms / 4 / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernels_ms": {k: round(v[1], 4) for k, v in kt.items()},
        "lookback": {"spin_limit": mgpu_hip.lookback_spin_limit(),
                     "fallbacks_in_timed_steps": fallbacks},
    }
    del keys_src, vals_src, keys, vals, aux
    torch.cuda.empty_cache()
    cc = copy_ceiling(torch, mgpu_hip)
    cc["scatter_frac_of_copy"] = achieved / cc["GBps"]
    result["roofline"]["copy_ceiling"] = cc
    log("copy ceiling", cc)
    torch.cuda.empty_cache()
    if not args.no_secondary:
        if not args.no_segsort:
            result["segsort"] = bench_segsort(torch, mgpu_hip, max(3, min(args.steps, 10)), 2,
                                             loop_baseline=not args.no_segsort_loop)
            torch.cuda.empty_cache()
        result["secondary"] = secondary(torch, mgpu_hip, max(3, min(args.steps, 10)), 2,
                                        chart=not args.no_chart)
        torch.cuda.empty_cache()
        if not args.no_vendor:
            try:  # a comparator; its failure must not cost the measurement above
                result["vendor_rocprim"] = bench_vendor(torch, mgpu_hip,
                                                        max(3, min(args.steps, 10)), 2)
            except (OSError, RuntimeError) as e:
                result["vendor_rocprim"] = {"error": str(e)}
                log("rocPRIM comparison failed:", e)
            torch.cuda.empty_cache()
        if not args.no_serving:
            result["serving"] = bench_serving(torch, mgpu_hip)
    if not args.no_cpu:
        log("cpu baseline 2^%d x %d" % (args.cpu_log2, args.cpu_reps))
        result["cpu_baseline"] = cpu_baseline(args.cpu_log2, args.cpu_reps)
        log("cpu baseline, threads")
        result["cpu_baseline_parallel"] = cpu_baseline_parallel(args.cpu_log2, args.cpu_reps)
        if not args.no_secondary:
            result["cpu_baselines_other"] = cpu_baselines_other()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
