"""Synthetic workloads of the fused-kernel benchmarks (shared by bench.py and
the tests).

Generator: SplitMix64 in counter form, seed 0x4859445241000000 + workload
number, so host and device produce bitwise identical inputs:
  integers in [0, 2^k):  word >> (64 - k)
  fp64 in [0, 1):        (word >> 11) * 2^-53
  fp64 in [-1, 1):       2*u - 1   (exact)
"""
import math

import numpy as np

SEED_BASE = 0x4859445241000000

# histogram: data[2^24] int64 uniform in [0, 2^20), 64 bins
HIST = dict(seed=SEED_BASE + 0, n=1 << 24, bits=20, bins=64)
# dot + map chain: xs, ys [2^27] fp64 uniform in [-1, 1); ys use counters n..2n-1
DOTMAP = dict(seed=SEED_BASE + 1, n=1 << 27)
# Jacobi heat: u[8192, 8192] fp64 uniform in [0, 1), 200 steps
JACOBI = dict(seed=SEED_BASE + 2, h=8192, w=8192, iters=200, alpha=0.1)
# batched histograms: 4096 instances
BATCH = dict(seed=SEED_BASE + 3, count=4096, bins=(64, 256, 4096, 65536), bits=31)
# served mix
SERVE = dict(seed=SEED_BASE + 4)

MASK64 = (1 << 64) - 1


def splitmix_at(seed, counter):
    """Pure-Python SplitMix64 word (counter form), for small host-side draws."""
    z = (seed + (counter + 1) * 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def unit(seed, counter):
    return (splitmix_at(seed, counter) >> 11) * 2.0 ** -53


def instance_seed(seed, k):
    """Per-instance data seed of the batched workload."""
    return (seed ^ ((k + 1) * 0xD1B54A32D192ED03)) & MASK64


def batch_instances(count=BATCH["count"], seed=BATCH["seed"], lo_log=10, hi_log=20,
                    bins=BATCH["bins"]):
    """Instance sizes n_k = floor(2^(lo + (hi-lo) u)) and bin counts drawn
    uniformly from `bins`.  Returns (lengths, bins) as int64 arrays."""
    lens = np.empty(count, np.int64)
    nb = np.empty(count, np.int64)
    for k in range(count):
        u = unit(seed, 2 * k)
        lens[k] = int(math.floor(2.0 ** (lo_log + (hi_log - lo_log) * u)))
        nb[k] = bins[splitmix_at(seed, 2 * k + 1) >> 62 if len(bins) == 4 else
                     splitmix_at(seed, 2 * k + 1) % len(bins)]
    return lens, nb


def batch_algorithmic_bytes(lens, bins):
    """Compulsory HBM traffic: read 8 B per element, write 8 B per bin."""
    return int(8 * np.sum(lens) + 8 * np.sum(bins))
