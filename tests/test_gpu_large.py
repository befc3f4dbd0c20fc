"""Inputs beyond 2^31 elements (64-bit indexing on the device, sizes the
reference's int64 loops allow), checked against the oracle on the same
SplitMix64 streams regenerated on the host chunk by chunk, so no 17 GB host
copy is needed.  Histogram and map chain bit-exact; dot within 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 31) + 5     # elements
CHUNK = 1 << 26       # host chunk


def chunks(n):
    for off in range(0, n, CHUNK):
        yield off, min(CHUNK, n - off)


def test_histogram_beyond_2_31(hyd, orc):
    seed = 0x4C41524745
    print("histogram n=%d" % N, flush=True)
    d = hyd.splitmix_i64(N, seed, bits=20)
    got = hyd.weighted_histogram(d, 64).numpy()
    d.free()
    ref = np.zeros(64, np.int64)
    with np.errstate(over="ignore"):
        for off, m in chunks(N):
            ref += orc.weighted_histogram(orc.splitmix_i64(m, seed, counter0=off, bits=20), 64)
    assert np.array_equal(got, ref)


def test_dot_beyond_2_31(hyd, orc):
    seed = 0x444F54
    print("dot n=%d" % N, flush=True)
    xs = hyd.splitmix_f64(N, seed, 0, symmetric=True)
    ys = hyd.splitmix_f64(N, seed, N, symmetric=True)
    got = float(hyd.dot(xs, ys).numpy()[0])
    xs.free()
    ys.free()
    ref = 0.0
    for off, m in chunks(N):
        ref += orc.dot_dd(orc.splitmix_f64(m, seed, off, symmetric=True),
                          orc.splitmix_f64(m, seed, N + off, symmetric=True))
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (got, ref)


def test_map_chain_beyond_2_31(hyd, orc):
    seed = 0x4D4150
    print("map chain n=%d" % N, flush=True)
    xs = hyd.splitmix_f64(N, seed, 0, symmetric=True)
    ys = hyd.splitmix_f64(N, seed, N, symmetric=True)
    out = hyd.map_chain(xs, ys)
    xs.free()
    ys.free()
    for off, m in chunks(N):
        if off not in (0, N - m) and (off // CHUNK) % 4:
            continue  # every 4th chunk plus both ends: the whole range of indices
        got = hyd.copy_out(out.ptr + 8 * off, m, np.float64)
        ref = orc.map_chain(orc.splitmix_f64(m, seed, off, symmetric=True),
                            orc.splitmix_f64(m, seed, N + off, symmetric=True))
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), off
    out.free()
