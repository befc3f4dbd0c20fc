"""GPU parity of the batched scatter (many histogram instances in one
persistent launch) against the oracle, for every policy, plus the autotuner."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def make_batch(orc, count, lo_log, hi_log, bins_set, seed):
    lens, bins = workloads.batch_instances(count=count, seed=seed, lo_log=lo_log, hi_log=hi_log,
                                           bins=bins_set)
    total = int(lens.sum())
    data = orc.splitmix_i64(total, seed + 1, bits=31)
    if total > 20:
        data[::17] = -data[::17]  # negative routes are dropped
    return data, lens, bins


@pytest.mark.parametrize("policy", [0, 1, 2])
@pytest.mark.parametrize("bins_set", [(64, 256, 4096, 65536), (1, 7, 100, 5000)])
def test_batched_matches_oracle(hyd, orc, policy, bins_set):
    print("batched policy=%d bins=%s" % (policy, bins_set), flush=True)
    data, lens, bins = make_batch(orc, 97, 0, 15, bins_set, 1000 + policy)
    off, hoff = hyd.batch_offsets(lens, bins)
    d = hyd.DeviceArray.from_numpy(data)
    got = hyd.batched_histogram(d, off, lens, bins, hoff, policy=policy).numpy()
    ref, rhoff = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(rhoff, hoff)
    assert np.array_equal(got, ref)


def test_batched_matches_single_calls(hyd, orc):
    data, lens, bins = make_batch(orc, 40, 8, 17, (64, 256, 4096, 65536), 77)
    off, hoff = hyd.batch_offsets(lens, bins)
    d = hyd.DeviceArray.from_numpy(data)
    got = hyd.batched_histogram(d, off, lens, bins, hoff).numpy()
    for k in range(lens.size):
        one = orc.weighted_histogram(data[off[k]:off[k] + lens[k]], int(bins[k]))
        assert np.array_equal(got[hoff[k]:hoff[k] + bins[k]], one), k


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_batched_chunks_and_bin_ranges(hyd, orc, policy):
    # long instances (several element chunks), bin counts beyond one LDS range
    # (16384) that are not a multiple of it, odd offsets (unaligned starts)
    lens = np.array([(1 << 18) + 5, 3 << 18, 100, 7, (1 << 18) + 1, 40_001], np.int64)
    bins = np.array([40_000, 64, 4096, 65536, 65536, 16_385], np.int64)
    off, hoff = hyd.batch_offsets(lens, bins)
    data = orc.splitmix_i64(int(lens.sum()), 4242, bits=31)
    data[::11] = -data[::11]
    d = hyd.DeviceArray.from_numpy(data)
    got = hyd.batched_histogram(d, off, lens, bins, hoff, policy=policy).numpy()
    ref, _ = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("policy", [0, 1])
def test_batched_split_pieces(hyd, orc, policy):
    # instances far longer than a queue's share with bins in one LDS range
    # (up to exactly 16384) are cut into pieces whose bins are added into the
    # zeroed output; pieces start at unaligned elements, and the batch mixes
    # them with short and multi-range instances
    lens = np.array([3_000_001, 17, 2_500_003, 1 << 21, 999_999, 5, 1_000_000], np.int64)
    bins = np.array([5000, 64, 16384, 1, 64, 3, 65536], np.int64)
    off, hoff = hyd.batch_offsets(lens, bins)
    data = orc.splitmix_i64(int(lens.sum()), 31337, bits=40)
    data[::13] = -data[::13]
    d = hyd.DeviceArray.from_numpy(data)
    print("batched split pieces policy=%d" % policy, flush=True)
    # the output starts as garbage and is reused: each call must overwrite it
    hist = hyd.DeviceArray.from_numpy(np.full(int(bins.sum()), 12345, np.int64))
    for _ in range(2):
        got = hyd.batched_histogram(d, off, lens, bins, hoff, hist=hist, policy=policy).numpy()
    ref, _ = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("policy", [0, 1])
def test_batched_split_multi_range(hyd, orc, policy):
    # instances with bins beyond one LDS range and far longer than a queue's
    # share are cut into pieces too (at most as many as fill the CUs with
    # their range tiles); power-of-two and other bin counts, unaligned piece
    # starts, negative elements, a garbage output reused twice
    lens = np.array([1 << 24, 3, 12_000_001, 50_000], np.int64)
    bins = np.array([65536, 64, 40_000, 65536], np.int64)
    off, hoff = hyd.batch_offsets(lens, bins)
    data = orc.splitmix_i64(int(lens.sum()), 5151, bits=40)
    data[::7] = -data[::7]
    d = hyd.DeviceArray.from_numpy(data)
    print("batched split multi-range policy=%d" % policy, flush=True)
    hist = hyd.DeviceArray.from_numpy(np.full(int(bins.sum()), -99, np.int64))
    for _ in range(2):
        got = hyd.batched_histogram(d, off, lens, bins, hoff, hist=hist, policy=policy).numpy()
    ref, _ = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_batched_wide_values(hyd, orc, policy):
    # 63-bit values, half negated: routes through both remainder forms (mask
    # for power-of-two bin counts, multiply-shift otherwise) over the full range
    lens = np.array([5000, 70_001, 3, 20_000, 1 << 17, 999], np.int64)
    bins = np.array([64, 1000, 7, 65536, 40_000, (1 << 20) + 3], np.int64)
    off, hoff = hyd.batch_offsets(lens, bins)
    data = orc.splitmix_i64(int(lens.sum()), 9191, bits=63)
    data[::2] = -data[::2]
    d = hyd.DeviceArray.from_numpy(data)
    got = hyd.batched_histogram(d, off, lens, bins, hoff, policy=policy).numpy()
    ref, _ = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(got, ref)


def test_batched_empty_and_zero_length(hyd, orc):
    lens = np.array([0, 5, 0, 1], np.int64)
    bins = np.array([3, 64, 1, 4096], np.int64)
    off, hoff = hyd.batch_offsets(lens, bins)
    data = np.arange(6, dtype=np.int64) + 3
    got = hyd.batched_histogram(hyd.DeviceArray.from_numpy(data), off, lens, bins, hoff).numpy()
    ref, _ = orc.batched_histograms(data, off, lens, bins)
    assert np.array_equal(got, ref)


def test_batched_full_workload(hyd, orc):
    c = workloads.BATCH
    lens, bins = workloads.batch_instances()
    off, hoff = hyd.batch_offsets(lens, bins)
    total = int(lens.sum())
    d = hyd.splitmix_i64(total, c["seed"], bits=c["bits"])
    got = hyd.batched_histogram(d, off, lens, bins, hoff).numpy()
    ref, _ = orc.batched_histograms(d.numpy(), off, lens, bins)
    assert np.array_equal(got, ref)


def test_autotuner_table_roundtrip(hyd, tmp_path):
    path = str(tmp_path / "tune.txt")
    n = hyd.tune_batched(path)
    assert n == 44  # 4 bin classes x 11 size classes
    table = hyd.tune_table()
    assert len(table) >= 44 and all(m in (0, 1) for _, _, m in table)
    text = open(path).read().split()
    assert len(text) == 3 * len(table)
