"""Helpers for the GPU parity tests: host <-> device copies and the case banner."""
import sys

import numpy as np
import torch

import mgpu_hip

DEV = "cuda:0"


def banner(*what):
    """Print and flush what names the case before its GPU step."""
    print("[case]", *what, flush=True)
    sys.stdout.flush()


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def to_host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty_like_np(a):
    return torch.empty(len(a), dtype=torch.from_numpy(a[:0].copy()).dtype, device=DEV)


def gen_dev(count, kind, seed=0, offset=0):
    dt = {"u32": torch.uint32, "lowentropy": torch.uint32, "i32_0_16": torch.int32,
          "f32_01": torch.float32, "u64": torch.uint64, "f32_pm1": torch.float32,
          "iota": torch.uint32}[kind]
    t = torch.empty(count, dtype=dt, device=DEV)
    mgpu_hip.generate(t, kind, seed, offset)
    return t


def as_signed_order(t):
    """int view whose signed order equals the unsigned order of t (u32/u64)."""
    if t.element_size() == 4:
        return t.view(torch.int32) ^ torch.tensor(-2 ** 31, dtype=torch.int32, device=t.device)
    return t.view(torch.int64) ^ torch.tensor(-2 ** 63, dtype=torch.int64, device=t.device)


def check_stable_sort_properties(keys_in, keys_out, vals_out):
    """Keys sorted, vals a permutation of 0..n-1, keys_out == keys_in[vals_out],
    and equal keys keep ascending vals: together these fix the stable sort of
    (key, original index) uniquely.  Unsigned keys, ascending."""
    n = keys_in.numel()
    ko = as_signed_order(keys_out)
    assert bool((ko[1:] >= ko[:-1]).all()), "keys not sorted"
    v = vals_out.view(torch.int32).long()
    seen = torch.zeros(n, dtype=torch.int32, device=keys_in.device)
    seen.index_add_(0, v, torch.ones(n, dtype=torch.int32, device=keys_in.device))
    assert bool((seen == 1).all()), "vals not a permutation"
    del seen
    gathered = as_signed_order(keys_in)[v]
    assert bool((gathered == ko).all()), "keys_out != keys_in[vals_out]"
    del gathered
    eq = ko[1:] == ko[:-1]
    vi = vals_out.view(torch.int32)
    assert bool(((vi[1:] > vi[:-1]) | ~eq).all()), "not stable"


def dev_equal(a, b):
    if a.element_size() == 4:
        return torch.equal(a.view(torch.int32), b.view(torch.int32))
    return torch.equal(a.view(torch.int64), b.view(torch.int64))
