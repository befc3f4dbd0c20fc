"""GPU parity of the guarded and chained scatters (scatter_guarded and
EScatterChain of src/Language/Hydrangea/Lowering.hs:1432-1517, 1622-1858)
against the oracle's serial loops: int64 (+) bit-exact under every
privatization strategy; fp64 (+.) bit-exact too, because the GPU reproduces
the serial per-bin order (stable sort by bin, in-order fold)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STRATS = [0, 1, 2, 3]  # auto, LDS, global, LDS split
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def dev(hyd, a):
    return hyd.DeviceArray.from_numpy(np.ascontiguousarray(a))


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def case_i64(n, ndst, seed):
    rng = np.random.default_rng(seed)
    idx = rng.integers(-2, ndst + 2, n)  # a few routes out of range: dropped
    vals = rng.integers(-2**62, 2**62, n)
    guard = rng.integers(0, 3, n)        # nonzero = true
    defaults = rng.integers(-2**40, 2**40, ndst)
    return idx, vals, guard, defaults


def test_guarded_known_answer(hyd, orc):
    (c,) = [c for c in json.load(open(GOLDEN)) if c["kind"] == "scatter_guarded"]
    i = c["inputs"]
    for dtype in (np.int64, np.float64):
        dst = dev(hyd, np.array(i["defaults"], dtype))
        hyd.scatter_guarded_add(dev(hyd, np.array(i["idx"], np.int64)), dev(hyd, np.array(i["vals"], dtype)),
                                dev(hyd, np.array(i["guard"], np.int64)), dst)
        assert dst.numpy().tolist() == [dtype(v) for v in c["expected"]]


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("n,ndst", [(1, 1), (7, 4), (1000, 64), (300_001, 64), (200_003, 5000),
                                    ((1 << 20) + 3, 40_000)])
def test_guarded_i64(hyd, orc, strategy, n, ndst):
    if strategy == 1 and ndst > 16383:
        pytest.skip("the LDS form needs the bins to fit LDS")
    print("guarded i64 n=%d ndst=%d strategy=%d" % (n, ndst, strategy), flush=True)
    idx, vals, guard, defaults = case_i64(n, ndst, n + ndst)
    dst = dev(hyd, defaults)
    hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals), dev(hyd, guard), dst, strategy)
    assert np.array_equal(dst.numpy(), orc.scatter_guarded_add(idx, vals, guard, defaults))


def test_guarded_i64_unguarded_and_misaligned(hyd, orc):
    idx, vals, guard, defaults = case_i64(50_001, 300, 3)
    dst = dev(hyd, defaults)
    hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals), None, dst)
    assert np.array_equal(dst.numpy(), orc.scatter_guarded_add(idx, vals, None, defaults))
    # 8-byte offset views take the scalar path
    di, dv, dg = dev(hyd, idx), dev(hyd, vals), dev(hyd, guard)
    dst = dev(hyd, defaults)
    hyd._check(hyd.lib().hydhip_scatter_guarded_add_i64(di.ptr + 8, dv.ptr + 8, dg.ptr + 8, idx.size - 1,
                                                        dst.ptr, 300, 0, None))
    assert np.array_equal(dst.numpy(), orc.scatter_guarded_add(idx[1:], vals[1:], guard[1:], defaults))


def test_guarded_all_false_and_empty(hyd, orc):
    idx, vals, _, defaults = case_i64(10_000, 64, 4)
    dst = dev(hyd, defaults)
    hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals), dev(hyd, np.zeros(idx.size, np.int64)), dst)
    assert np.array_equal(dst.numpy(), defaults)
    dst = dev(hyd, defaults.astype(np.float64))
    hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals.astype(np.float64)),
                            dev(hyd, np.zeros(idx.size, np.int64)), dst)
    assert bits(dst.numpy()).tolist() == bits(defaults.astype(np.float64)).tolist()
    e = hyd.DeviceArray((0,), np.int64)
    assert hyd.lib().hydhip_scatter_guarded_add_f64(e.ptr, e.ptr, e.ptr, 0, dst.ptr, 64, None) == 0


@pytest.mark.parametrize("n,ndst", [(6, 2), (3000, 7), (262_147, 11), (262_147, 100_000), (1 << 20, 1 << 20)])
def test_guarded_f64_serial_order(hyd, orc, n, ndst):
    print("guarded f64 n=%d ndst=%d" % (n, ndst), flush=True)
    rng = np.random.default_rng(n + ndst)
    if n == 6:
        idx = np.array([0, 0, 0, 1, 1, 1])
        vals = np.array([1e16, 1.0, -1e16, 1e16, -1e16, 1.0])
        guard = np.ones(6, np.int64)
        defaults = np.zeros(2)
    else:
        idx = rng.integers(-1, ndst + 1, n)
        vals = rng.standard_normal(n) * 10.0 ** rng.integers(-12, 13, n)
        guard = rng.integers(0, 2, n)
        defaults = rng.standard_normal(ndst)
    dst = dev(hyd, defaults)
    hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals), dev(hyd, guard), dst)
    ref = orc.scatter_guarded_add(idx, vals, guard, defaults)
    assert np.array_equal(bits(dst.numpy()), bits(ref))


def long_run_cases():
    """Per-bin (old value, run) pairs for the long-run fold: each run is longer
    than one thread's share, so a wave folds it, and the cases sit on both
    sides of the exact-sum shortcut (every operand a multiple of 2^k and the
    magnitudes summing below 2^(53+k))."""
    rng = np.random.default_rng(21)
    one = np.ones
    cases = [
        (0.25, rng.integers(-2**20, 2**20, 5000) / 64.0),                 # dyadic: exact
        (0.0, np.concatenate([[2.0**53], one(100)])),                      # S >= 2^53: 2^53 + 1 + ... ties to even
        (2.0**53 - 64, one(63)),                                           # S = 2^53 - 1: exact
        (2.0**53 - 64, one(64)),                                           # S = 2^53: serial, exact anyway
        (2.0**53 - 64, np.concatenate([one(63), [0.5]])),                  # k = -1: the last add ties to even
        (-0.0, np.full(40, -0.0)),                                         # -0 kept
        (0.0, np.full(40, -0.0)),                                          # +0 + -0 = +0
        (-0.0, np.concatenate([np.full(39, -0.0), [0.0]])),
        (1.0, np.concatenate([one(20), [np.nan], one(20)])),
        (1.0, np.concatenate([one(20), [np.inf], one(20)])),
        (0.0, np.concatenate([[np.inf], one(20), [-np.inf], one(20)])),
        (0.0, rng.integers(0, 2**40, 100) * 2.0**-1074),                   # subnormal multiples
        (0.0, np.full(40, 2.0**971)),                                      # k = 971: serial
        (0.0, np.full(40, 1e308)),                                         # overflows to inf
        (3.0, rng.standard_normal(3000)),                                  # no shortcut
        (0.0, np.concatenate([[1e16], one(33), [-1e16]])),                 # absorbed ones
    ]
    return cases


def test_guarded_f64_long_runs(hyd, orc):
    # The wave-per-run fold, both sides of its exact-sum shortcut, against the
    # serial loop bit for bit; bins interleaved in element order.
    cases = long_run_cases()
    ndst = len(cases) + 3
    rng = np.random.default_rng(5)
    labels = np.concatenate([np.full(len(v), b) for b, (_, v) in enumerate(cases)])
    rng.shuffle(labels)
    vals = np.empty(labels.size)
    for b, (_, v) in enumerate(cases):
        vals[labels == b] = v               # each bin keeps its run's order
    defaults = np.concatenate([[o for o, _ in cases], [7.0, -7.0, 0.5]])
    guard = np.ones(labels.size, np.int64)
    print("guarded f64 long runs n=%d bins=%d" % (labels.size, ndst), flush=True)
    dst = dev(hyd, defaults)
    hyd.scatter_guarded_add(dev(hyd, labels.astype(np.int64)), dev(hyd, vals), dev(hyd, guard), dst)
    ref = orc.scatter_guarded_add(labels.astype(np.int64), vals, guard, defaults)
    got = dst.numpy()
    same = (bits(got) == bits(ref)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), [(b, got[b], ref[b]) for b in np.flatnonzero(~same)]
    # the cases do separate the two orders: the serial loop loses the ones
    assert ref[1] == 2.0**53 and ref[4] == 2.0**53 and ref[15] == 0.0
    assert np.signbit(ref[5]) and not np.signbit(ref[6]) and not np.signbit(ref[7])


def test_chain_i64(hyd, orc):
    rng = np.random.default_rng(8)
    ndst = 3000
    defaults = rng.integers(-100, 100, ndst)
    phases, ref = [], defaults.copy()
    for k, m in enumerate((100_000, 0, 7, 250_001)):
        idx, vals, guard, _ = case_i64(m, ndst, 50 + k)
        g = None if k == 2 else guard
        phases.append((idx, vals, g))
        ref = orc.scatter_guarded_add(idx, vals, g, ref)
    dst = dev(hyd, defaults)
    hyd.scatter_chain_add([(dev(hyd, i), dev(hyd, v), None if g is None else dev(hyd, g)) for i, v, g in phases],
                          dst)
    assert np.array_equal(dst.numpy(), ref)


def test_chain_f64_phase_order(hyd, orc):
    rng = np.random.default_rng(9)
    ndst = 13
    phases = []
    for k, m in enumerate((5000, 0, 3001, 70_000)):
        phases.append((rng.integers(0, ndst, m), rng.standard_normal(m) * 10.0 ** rng.integers(-9, 10, m),
                       None if k == 0 else rng.integers(0, 2, m)))
    defaults = rng.standard_normal(ndst)
    dst = dev(hyd, defaults)
    hyd.scatter_chain_add([(dev(hyd, i), dev(hyd, v), None if g is None else dev(hyd, g)) for i, v, g in phases],
                          dst)
    ref = orc.scatter_chain_add_f64(phases, defaults)
    assert np.array_equal(bits(dst.numpy()), bits(ref))


def test_voxel_benchmarks_full_size(hyd, orc):
    # the bench workloads themselves, bit-exact against the serial loops
    import workloads
    for gen, cfg, dtype in ((workloads.voxel_rasterization, workloads.VOXEL, np.int64),
                            (workloads.voxel_trilinear_splat, workloads.SPLAT, np.float64)):
        idx, vals, guard, nvox = gen(**cfg)
        vals = vals.astype(dtype)
        print("voxel %s points=%d contributions=%d" % (dtype.__name__, cfg["points"], idx.size), flush=True)
        dst = dev(hyd, np.zeros(nvox, dtype))
        hyd.scatter_guarded_add(dev(hyd, idx), dev(hyd, vals), dev(hyd, guard), dst)
        ref = orc.scatter_guarded_add(idx, vals, guard, np.zeros(nvox, dtype))
        assert np.array_equal(bits(dst.numpy()), bits(ref))
