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


def serve_mix(count, seed=SERVE["seed"], max_side_log=12, max_hist_log=24, max_dot_log=26,
              nnz_lo=1e5, nnz_hi=1e7):
    """The served mix: 40% Jacobi stencils (side floor(2^(8+4u)), 10-100
    steps), 30% 64-bin histograms (n = floor(2^(12+12u))), 20% fp64 dots
    (n = floor(2^(14+12u))), 10% COO->CSR builds over 2^20 x 2^20
    (nnz log-uniform in [1e5, 1e7]).  Upper ends can be lowered for tests."""
    reqs = []
    for k in range(count):
        c = 4 * k
        u = unit(seed, c)
        v = unit(seed, c + 1)
        w = unit(seed, c + 2)
        if u < 0.4:
            side = int(math.floor(2.0 ** (8 + (max_side_log - 8) * v)))
            steps = 10 + int(w * 91) if w < 1 else 100
            reqs.append(dict(kind="stencil", side=side, iters=min(steps, 100)))
        elif u < 0.7:
            reqs.append(dict(kind="hist", n=int(math.floor(2.0 ** (12 + (max_hist_log - 12) * v))), bins=64))
        elif u < 0.9:
            reqs.append(dict(kind="dot", n=int(math.floor(2.0 ** (14 + (max_dot_log - 14) * v)))))
        else:
            nnz = int(nnz_lo * (nnz_hi / nnz_lo) ** v)
            reqs.append(dict(kind="coo", nnz=nnz, nrows=1 << 20, ncols=1 << 20))
    return reqs


def request_bytes_algorithmic(r):
    """Compulsory HBM bytes of one served request."""
    if r["kind"] == "stencil":
        return 16 * r["side"] ** 2 * r["iters"]
    if r["kind"] == "hist":
        return 8 * r["n"] + 8 * r["bins"]
    if r["kind"] == "dot":
        return 16 * r["n"]
    return 40 * r["nnz"] + 8 * (r["nrows"] + 1)


def batch_algorithmic_bytes(lens, bins):
    """Compulsory HBM traffic: read 8 B per element, write 8 B per bin."""
    return int(8 * np.sum(lens) + 8 * np.sum(bins))


# Guarded scatters: the reference's voxel benchmarks (bench/voxel_rasterization,
# bench/voxel_trilinear_splat) with their default grid 64^3 and keep period 3,
# their route/weight/keep functions materialised as arrays.
VOXEL = dict(points=5_000_000, nx=64, ny=64, nz=64, keep_period=3)
SPLAT = dict(points=1_000_000, nx=64, ny=64, nz=64, keep_period=3)


def voxel_rasterization(points, nx, ny, nz, keep_period):
    """route/weight/keep of bench/voxel_rasterization/voxel_rasterization.hyd:
    imod x m = x - (x / m) * m; all operands non-negative, so C division is
    floor division.  Returns (idx, vals, guard) int64 and nvox."""
    p = np.arange(points, dtype=np.int64)
    x = (p * 17 + 3) % nx
    y = (p * 29 + 5) % ny
    z = (p * 43 + 7) % nz
    route = ((z * ny) + y) * nx + x
    weight = ((x + 1) * (y + 2)) + (z * 3) + 1
    keep = ((p // keep_period) * keep_period == p).astype(np.int64)
    return route, weight, keep, nx * ny * nz


def voxel_trilinear_splat(points, nx, ny, nz, keep_period):
    """route_of/weight_of/keep_of of bench/voxel_trilinear_splat/
    voxel_trilinear_splat.hyd over 8 corner contributions per point; the
    weight is float_of(base * wx * wy * wz) /. 64.0 (exact in fp64 here).
    Returns (idx int64, vals fp64, guard int64) and nvox."""
    i = np.arange(points * 8, dtype=np.int64)
    p = i // 8
    c = i % 8
    bx, by, bz = c % 2, (c // 2) % 2, (c // 4) % 2
    x = (p * 17 + 3 + bx) % nx
    y = (p * 29 + 5 + by) % ny
    z = (p * 43 + 7 + bz) % nz
    route = ((z * ny) + y) * nx + x
    fx, fy, fz = (p * 5 + 1) % 4, (p * 7 + 2) % 4, (p * 11 + 3) % 4
    wx = (1 - bx) * (4 - fx) + bx * fx
    wy = (1 - by) * (4 - fy) + by * fy
    wz = (1 - bz) * (4 - fz) + bz * fz
    base = (p * 3) + 1
    weight = (base * wx * wy * wz).astype(np.float64) / 64.0
    keep = ((p // keep_period) * keep_period == p).astype(np.int64)
    return route, weight, keep, nx * ny * nz
