"""GPU parity at the stated workload scales (the bench's own shapes):
  * key-only stable mergesort, less, of 2^20 uint32 keys uniform from seed 0,
    against std::stable_sort;
  * batched segmented sort of 65536 lognormal segments (median 256, sigma
    1.5, clamped to [1, 2^20]) of uint32 key-value pairs from seed 0, with the
    autotuned table, recorded once into a graph and replayed, against one
    std::stable_sort per segment (every bucket, the persistent kernels'
    stride loop and the large-segment radix path run);
  * the serving engine on sort / merge / scan requests of u32 / u64 / f32
    with sizes log-uniform over 2^10..2^26 and a 64 GiB budget, saturated and
    then under Poisson arrivals at 0.8x the measured saturation rate, every
    response verified."""
import numpy as np
import pytest
import torch

import mgpu_hip
import oracle
from gpu_util import banner, to_dev, to_host
from mgpu_hip import serving, workloads

pytestmark = pytest.mark.gpu


def test_mergesort_keys_u32_2p20_seed0():
    n = 1 << 20
    banner("mergesort keys u32 2^20 seed 0")
    keys = oracle.generate(n, "u32", 0)
    dk = to_dev(keys)
    mgpu_hip.mergesort_keys(dk, "less")
    np.testing.assert_array_equal(to_host(dk), oracle.stable_sort(keys))


def test_segmented_sort_65536_lognormal_graph_replays():
    lens = workloads.lognormal_lengths(65536, 256.0, 1.5, 1, 1 << 20, seed=0)
    starts = workloads.segment_starts(lens)
    n = int(lens.sum())
    banner("segsort 65536 lognormal kv u32, autotuned, graph", n)
    keys = oracle.generate(n, "u32", 0)
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = oracle.segmented_stable_sort(keys, starts, vals, "u32")
    tuning = mgpu_hip.segmented_sort_autotune("u32", True, "less")
    ks, vs = to_dev(keys), to_dev(vals)
    segs = to_dev(starts)
    k, v = ks.clone(), vs.clone()
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        aux = mgpu_hip.segmented_sort(k, segs, v, tuning=tuning)  # eager, sizes the aux
        stream.synchronize()
        assert np.array_equal(k.cpu().numpy(), ek) and np.array_equal(v.cpu().numpy(), ev)
        g = mgpu_hip.Graph(stream)
        with g:
            mgpu_hip.segmented_sort(k, segs, v, tuning=tuning, aux=aux, stream=stream)
        for rep in range(2):
            k.copy_(ks)
            v.copy_(vs)
            g.launch()
            stream.synchronize()
            assert np.array_equal(k.cpu().numpy(), ek), "keys differ on replay %d" % rep
            assert np.array_equal(v.cpu().numpy(), ev), "values differ on replay %d" % rep
    torch.cuda.synchronize()


def _run_and_verify(reqs, rate, budget):
    sv = serving.Server(slots=8, budget_bytes=budget)
    try:
        resps = serving.run_load(sv, reqs, rate=rate)
        assert len(resps) == len(reqs)
        summ = serving.summarize(resps, reqs)
        for r in resps:
            serving.verify(r)
            sv.release(r["id"])
        st = sv.stats()
        assert st["in_use"] == 0 and st["pending"] == 0 and st["in_flight"] == 0
        assert st["peak"] <= budget
    finally:
        sv.close()
    torch.cuda.synchronize()
    return summ


def test_serving_full_size_mix_saturated_and_poisson():
    budget = 64 << 30
    banner("serving 2^10..2^26 mix, saturation")
    reqs = serving.make_requests(64, seed=11, lo_log2=10, hi_log2=26)
    assert max(serving.elements(r) for r in reqs) > 1 << 24
    sat = _run_and_verify(reqs, None, budget)
    rate = 0.8 * sat["requests_per_s"]
    banner("serving 2^10..2^26 mix, Poisson at 0.8x saturation", rate)
    summ = _run_and_verify(serving.make_requests(64, seed=12, lo_log2=10, hi_log2=26), rate,
                           budget)
    print("saturation", sat, "poisson", summ, flush=True)
    assert summ["p99_ms"] >= summ["p50_ms"] > 0
