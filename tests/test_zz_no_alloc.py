"""hydhip_weighted_histogram_i64_noalloc (the entry point that allocates no
device memory) against hydhip_weighted_histogram_i64 on the same inputs,
through the C ABI: bit-exact, on two calls in a row, the second with a
smaller input, for every strategy and bin class; and the oracle agrees."""
import ctypes

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

MAX_LDS_BINS = 8192  # the no-allocation scratch: 8192 words


def call(hyd, name, ptr, n, nbins, strategy):
    out = hyd.DeviceArray((nbins,), np.int64)
    L = hyd.lib()
    rc = getattr(L, name)(ptr, n, nbins, out.ptr, strategy, None)
    if rc != hyd.HYD_OK:
        raise hyd.HydError(rc, L.hydhip_last_error().decode())
    return out.numpy()


@pytest.mark.parametrize("nbins", [1, 7, 64, 1000, 4096, 8192, 65536])
@pytest.mark.parametrize("strategy", [0, 1, 2, 3])
def test_noalloc_equals_existing_two_calls(hyd, orc, nbins, strategy):
    if strategy == 1 and nbins > MAX_LDS_BINS:
        pytest.skip("the LDS form beyond the scratch is rejected (tested below)")
    # call 1, then call 2 on a smaller input (fewer elements and bins): the
    # second must not see what the first left in the shared scratch
    for n, nb in ((1 << 20) + 3, nbins), (50_001, max(1, nbins // 4)):
        print("noalloc n=%d bins=%d strategy=%d" % (n, nb, strategy), flush=True)
        data = orc.splitmix_i64(n, 0x4E41 + n + nb, bits=31)
        data[::11] = -data[::11]  # negative routes: dropped
        d = hyd.DeviceArray.from_numpy(data)
        got = call(hyd, "hydhip_weighted_histogram_i64_noalloc", d.ptr, n, nb, strategy)
        ref = call(hyd, "hydhip_weighted_histogram_i64", d.ptr, n, nb, strategy)
        assert np.array_equal(got, ref), (n, nb, strategy)
        assert np.array_equal(got, orc.weighted_histogram(data, nb))
        d.free()


def test_noalloc_misaligned_empty_and_full_size(hyd, orc):
    n = 100_001
    data = orc.splitmix_i64(n + 1, 0x4E42, bits=20)
    d = hyd.DeviceArray.from_numpy(data)
    for strategy in (0, 1, 2, 3):
        # an 8-byte offset takes the scalar-load path
        got = call(hyd, "hydhip_weighted_histogram_i64_noalloc", d.ptr + 8, n, 64, strategy)
        assert np.array_equal(got, call(hyd, "hydhip_weighted_histogram_i64", d.ptr + 8, n, 64, strategy))
        assert np.array_equal(got, orc.weighted_histogram(data[1:], 64))
    # n = 0: fill [nbins] 0
    assert not call(hyd, "hydhip_weighted_histogram_i64_noalloc", d.ptr, 0, 64, 0).any()
    d.free()
    c = workloads.HIST
    full = hyd.splitmix_i64(c["n"], c["seed"], bits=c["bits"])
    got = call(hyd, "hydhip_weighted_histogram_i64_noalloc", full.ptr, c["n"], c["bins"], 0)
    assert np.array_equal(got, call(hyd, "hydhip_weighted_histogram_i64", full.ptr, c["n"], c["bins"], 0))
    assert np.array_equal(got, orc.weighted_histogram(full.numpy(), c["bins"]))
    full.free()


def test_noalloc_rejects_lds_beyond_scratch(hyd, orc):
    n = 4096
    d = hyd.DeviceArray.from_numpy(orc.splitmix_i64(n, 0x4E43, bits=20))
    out = hyd.DeviceArray((MAX_LDS_BINS + 1,), np.int64)
    L = hyd.lib()
    rc = L.hydhip_weighted_histogram_i64_noalloc(d.ptr, n, MAX_LDS_BINS + 1, out.ptr, 1, None)
    assert rc == hyd.HYD_E_INVALID
    assert b"no-allocation scratch" in L.hydhip_last_error()
    # the same size with the automatic choice takes a form without scratch
    got = call(hyd, "hydhip_weighted_histogram_i64_noalloc", d.ptr, n, MAX_LDS_BINS + 1, 0)
    assert np.array_equal(got, orc.weighted_histogram(d.numpy(), MAX_LDS_BINS + 1))
    d.free()
