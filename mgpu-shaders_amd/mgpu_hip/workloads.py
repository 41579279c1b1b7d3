"""Deterministic synthetic workloads (host side, numpy).

Counter-based SplitMix64 (the same mixing as the device generator):
  * lognormal segment lengths: Box-Muller normals from two uniforms,
    round(exp(ln(median) + sigma * z)) clamped to [lo, hi];
  * Poisson arrivals: inter-arrival -ln(1 - U) / rate;
  * log-uniform sizes: round(2 ** (lo + (hi - lo) * U)).
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix_u64(seed, stream, n):
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(40)) + (np.uint64(stream) << np.uint64(36)) + i + _G
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, stream, n):
    """Doubles in [0, 1) with 53 random bits."""
    return (splitmix_u64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def lognormal_lengths(num_segments, median=256.0, sigma=1.5, lo=1, hi=1 << 20, seed=0):
    u1 = uniform01(seed, 0, num_segments)
    u2 = uniform01(seed, 1, num_segments)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    lens = np.rint(np.exp(np.log(median) + sigma * z))
    return np.clip(lens, lo, hi).astype(np.int64)


def segment_starts(lengths):
    """int32 starts of segments with the given lengths (first = 0)."""
    starts = np.zeros(len(lengths), dtype=np.int64)
    if len(lengths) > 1:
        starts[1:] = np.cumsum(lengths[:-1])
    return starts.astype(np.int32)


def poisson_arrivals(n, rate, seed=0):
    """Arrival times (seconds) of n requests at `rate` per second."""
    gaps = -np.log1p(-uniform01(seed, 2, n)) / rate
    return np.cumsum(gaps)


def loguniform_sizes(n, lo_log2=10, hi_log2=26, seed=0):
    u = uniform01(seed, 3, n)
    return np.rint(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * u)).astype(np.int64)
