"""GPU parity: every libhydhip kernel against the CPU oracle on the same
seeded inputs.  Integer/index results and fp64 elementwise/stencil results
must be bit-exact; fp64 reductions within 1e-12 relative (the bound stated
for fp64 reductions)."""
import json
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
REL_TOL_REDUCTION = 1e-12


def golden(kind):
    with open(GOLDEN) as f:
        return [c for c in json.load(f) if c["kind"] == kind]


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float64:
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return a.shape == b.shape and np.array_equal(a, b)


def dev(hyd, a):
    return hyd.DeviceArray.from_numpy(np.ascontiguousarray(a))


# ---- generator ------------------------------------------------------------------

def test_generators_match_oracle(hyd, orc):
    n = 1_000_003
    s = workloads.HIST["seed"]
    assert bits_equal(hyd.splitmix_i64(n, s, 3, bits=20).numpy(), orc.splitmix_i64(n, s, 3, bits=20))
    assert bits_equal(hyd.splitmix_f64(n, s, 9, symmetric=True).numpy(),
                      orc.splitmix_f64(n, s, 9, symmetric=True))


# ---- histogram / scatter ------------------------------------------------------------

STRATS = [0, 1, 2, 3]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 1001, (1 << 20) + 3])
@pytest.mark.parametrize("nbins", [1, 64, 256, 4096, 65536])
@pytest.mark.parametrize("strategy", STRATS)
def test_weighted_histogram(hyd, orc, n, nbins, strategy):
    if strategy == 1 and nbins > 16384:
        pytest.skip("LDS form needs the bins to fit LDS")
    print("hist n=%d bins=%d strategy=%d" % (n, nbins, strategy), flush=True)
    data = orc.splitmix_i64(n, 0x5151 + n + nbins, bits=31)
    if n > 10:
        data[::13] = -data[::13]  # negative routes: dropped
    got = hyd.weighted_histogram(dev(hyd, data), nbins, strategy).numpy()
    assert bits_equal(got, orc.weighted_histogram(data, nbins))


@pytest.mark.parametrize("nbins", [3, 64, 1000, 65536])
def test_weighted_histogram_wide_values(hyd, orc, nbins):
    # 63-bit and negative values: the 64-bit remainder path and wrapping squares
    data = orc.splitmix_i64(300_001, 0x77 + nbins, bits=63)
    data[::3] = -data[::3]
    data[::5] = data[::5] >> 40
    for strategy in STRATS:
        if strategy == 1 and nbins > 16384:
            continue
        got = hyd.weighted_histogram(dev(hyd, data), nbins, strategy).numpy()
        assert bits_equal(got, orc.weighted_histogram(data, nbins)), strategy


def test_weighted_histogram_misaligned(hyd, orc):
    # an 8-byte offset view takes the scalar-load path
    n = 100_001
    data = orc.splitmix_i64(n + 1, 99, bits=20)
    d = dev(hyd, data)
    out = hyd.DeviceArray((64,), np.int64)
    for strategy in STRATS:
        hyd._check(hyd.lib().hydhip_weighted_histogram_i64(d.ptr + 8, n, 64, out.ptr, strategy, None))
        assert bits_equal(out.numpy(), orc.weighted_histogram(data[1:], 64))


def test_weighted_histogram_full_size(hyd, orc):
    c = workloads.HIST
    d = hyd.splitmix_i64(c["n"], c["seed"], bits=c["bits"])
    host = d.numpy()
    ref = orc.weighted_histogram(host, c["bins"])
    for strategy in STRATS:
        got = hyd.weighted_histogram(d, c["bins"], strategy).numpy()
        assert bits_equal(got, ref), strategy
    # closed-form checksum: sum of bins == sum of squares (mod 2^64)
    sq = (host.astype(np.uint64) ** 2).sum(dtype=np.uint64)
    assert ref.astype(np.uint64).sum(dtype=np.uint64) == sq


@pytest.mark.parametrize("case", golden("scatter"), ids=[c["name"] for c in golden("scatter")])
@pytest.mark.parametrize("strategy", STRATS)
def test_scatter_known_answers(hyd, case, strategy):
    i = case["inputs"]
    dst = dev(hyd, np.zeros(i["ndst"], np.int64))
    hyd.scatter_add(dev(hyd, np.array(i["idx"], np.int64)), dev(hyd, np.array(i["vals"], np.int64)), dst,
                    strategy)
    assert dst.numpy().tolist() == case["expected"]


def test_weighted_histogram_closed_form(hyd):
    (c,) = golden("weighted_histogram")
    got = hyd.weighted_histogram(dev(hyd, np.array(c["inputs"]["data"], np.int64)), 64).numpy()
    assert got.tolist() == c["expected"]


@pytest.mark.parametrize("strategy", STRATS)
def test_scatter_add_accumulates(hyd, orc, strategy):
    n, ndst = 300_007, 5000
    idx = orc.splitmix_i64(n, 7, bits=40) % ndst
    vals = orc.splitmix_i64(n, 8)  # full 64-bit values: wraparound
    base = orc.splitmix_i64(ndst, 9)
    dst = dev(hyd, base)
    hyd.scatter_add(dev(hyd, idx), dev(hyd, vals), dst, strategy)
    assert bits_equal(dst.numpy(), orc.scatter_add(idx, vals, base))


# ---- dot ------------------------------------------------------------------------------

@pytest.mark.parametrize("case", golden("dot"), ids=[c["name"] for c in golden("dot")])
def test_dot_known(hyd, case):
    i = case["inputs"]
    got = hyd.dot(dev(hyd, np.array(i["xs"])), dev(hyd, np.array(i["ys"]))).numpy()[0]
    assert got == case["expected"]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 257, 1_000_003])
def test_dot_random(hyd, orc, n):
    xs = orc.splitmix_f64(n, 101 + n, symmetric=True)
    ys = orc.splitmix_f64(n, 101 + n, counter0=n, symmetric=True)
    dx, dy = dev(hyd, xs), dev(hyd, ys)
    got = hyd.dot(dx, dy, init=0.5).numpy()[0]
    ref = orc.dot(xs, ys, init=0.5)
    exact = orc.dot_dd(xs, ys, init=0.5)
    assert abs(got - ref) <= REL_TOL_REDUCTION * abs(ref)
    assert abs(got - exact) <= REL_TOL_REDUCTION * abs(exact)
    # deterministic: same bits on a second run
    assert hyd.dot(dx, dy, init=0.5).numpy()[0] == got


def test_dot_full_size(hyd, orc):
    c = workloads.DOTMAP
    n = c["n"]
    xs = hyd.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = hyd.splitmix_f64(n, c["seed"], n, symmetric=True)
    got = hyd.dot(xs, ys).numpy()[0]
    hx, hy = xs.numpy(), ys.numpy()
    ref = orc.dot(hx, hy)
    assert abs(got - ref) <= REL_TOL_REDUCTION * abs(ref), (got, ref)
    one = hyd.DeviceArray((n,), np.float64)
    hyd._check(hyd.lib().hydhip_fill_splitmix_f64(one.ptr, n, 0, 0, 0.0, 1.0, None))  # all 1.0
    assert hyd.dot(one, one).numpy()[0] == float(n)


# ---- map chain -------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 1001, 1_000_003])
def test_map_chain_random(hyd, orc, n):
    xs = orc.splitmix_f64(n, 201 + n, symmetric=True)
    ys = orc.splitmix_f64(n, 201 + n, counter0=n, symmetric=True)
    got = hyd.map_chain(dev(hyd, xs), dev(hyd, ys)).numpy()
    assert bits_equal(got, orc.map_chain(xs, ys))


def test_map_chain_misaligned_and_special(hyd, orc):
    xs = np.array([7.0, 0.0, -0.0, 1.0, -1.0, np.nan, 0.25, 1e308, -1e308, np.inf, -np.inf, 3e5])
    ys = np.array([7.0, -0.0, 0.0, np.nan, 2.0, 1.0, -0.5, 1e308, 1.0, 1.0, -np.inf, -1e-300])
    dx, dy = dev(hyd, xs), dev(hyd, ys)
    out = hyd.DeviceArray((xs.size - 1,), np.float64)
    hyd._check(hyd.lib().hydhip_map_chain_f64(dx.ptr + 8, dy.ptr + 8, xs.size - 1, out.ptr, None))
    ref = orc.map_chain(xs[1:], ys[1:])
    got = out.numpy()
    assert np.array_equal(got, ref, equal_nan=True)
    assert bits_equal(got[~np.isnan(got)], ref[~np.isnan(ref)])


def test_map_chain_full_size(hyd, orc):
    c = workloads.DOTMAP
    n = c["n"]
    xs = hyd.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = hyd.splitmix_f64(n, c["seed"], n, symmetric=True)
    got = hyd.map_chain(xs, ys).numpy()
    ref = orc.map_chain(xs.numpy(), ys.numpy())
    assert bits_equal(got, ref)


def test_math_exp_sqrt_match_libm(hyd, orc):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-760, 760, 2_000_000), rng.uniform(0, 0.003, 2_000_000),
                        np.frombuffer(rng.integers(0, 2**63, 1_000_000, dtype=np.uint64).tobytes(),
                                      np.float64),
                        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 709.78, 709.79, -745.13, -745.14])])
    dx = dev(hyd, x)
    out = hyd.DeviceArray(x.shape, np.float64)
    hyd._check(hyd.lib().hydhip_math_f64(0, dx.ptr, x.size, out.ptr, None))
    got, ref = out.numpy(), orc.libm_exp(x)
    same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), x[~same][:5]
    ax = np.abs(x)
    dax = dev(hyd, ax)
    hyd._check(hyd.lib().hydhip_math_f64(1, dax.ptr, x.size, out.ptr, None))
    got, ref = out.numpy(), orc.libm_sqrt(ax)
    same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), ax[~same][:5]


# ---- stencils -------------------------------------------------------------------------------

SHAPES = [(1, 1), (1, 9), (7, 1), (5, 6), (67, 130), (64, 64), (65, 129), (300, 1000), (1000, 63)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("body", [0, 1, 2, 3])
def test_stencil2d(hyd, orc, shape, body):
    u = orc.splitmix_f64(shape[0] * shape[1], 301 + body + shape[0]).reshape(shape)
    got = hyd.stencil2d(dev(hyd, u), body, 0.1).numpy()
    assert bits_equal(got, orc.stencil2d(u, body, 0.1))


def test_stencil_known_answers(hyd):
    g = np.arange(1, 10, dtype=np.float64).reshape(3, 3)
    assert hyd.stencil2d(dev(hyd, g), hyd.STENCIL_JACOBI4).numpy()[0, 0] * 4.0 == 8.0
    assert hyd.stencil2d(dev(hyd, g), hyd.STENCIL_LAPLACIAN).numpy()[1, 1] == 0.0
    c = np.full((33, 70), 0.37)
    assert bits_equal(hyd.stencil2d(dev(hyd, c), hyd.STENCIL_HEAT, 0.1).numpy(), c)


# the temporal kernel's output tile is 48 rows x 240 columns (a 64 x 256 box,
# 4 columns per lane): widths at tile edges, and 489 where the second tile
# column is the first interior one (its box ends at column 488 = w - 1)
TEMPORAL_EDGES = [(97, 240), (49, 241), (130, 481), (300, 489), (113, 720)]


@pytest.mark.parametrize("shape", SHAPES + [(49, 97), (96, 96), (200, 47)] + TEMPORAL_EDGES)
@pytest.mark.parametrize("body", [0, 1, 2, 3])
def test_iterate_temporal_blocks(hyd, orc, shape, body):
    # 11 steps = one fused pass of 8 and one of 3, on tile-unaligned shapes
    u = orc.splitmix_f64(shape[0] * shape[1], 351 + body + shape[1]).reshape(shape)
    res, _, _ = hyd.iterate_stencil2d(dev(hyd, u), 11, body=body, alpha=0.1, use_graph=False)
    assert bits_equal(res.numpy(), orc.iterate_stencil2d(u, 11, body, 0.1))


@pytest.mark.parametrize("iters", [0, 1, 2, 5, 8, 9, 17])
@pytest.mark.parametrize("graph", [False, True])
def test_iterate(hyd, orc, iters, graph):
    u = orc.splitmix_f64(70 * 135, 401 + iters).reshape(70, 135)
    res, b0, b1 = hyd.iterate_stencil2d(dev(hyd, u), iters, use_graph=graph)
    assert bits_equal(res.numpy(), orc.iterate_stencil2d(u, iters))


def test_iterate_graph_replays(hyd, orc):
    # the cached graph replayed several times, with eager work in between
    u = orc.splitmix_f64(130 * 70, 409).reshape(130, 70)
    du = dev(hyd, u)
    ref = orc.iterate_stencil2d(u, 6)
    b0, b1 = hyd.DeviceArray(u.shape, np.float64), hyd.DeviceArray(u.shape, np.float64)
    for _ in range(3):
        res, _, _ = hyd.iterate_stencil2d(du, 6, buf0=b0, buf1=b1, use_graph=True)
        hyd.sync()
        assert bits_equal(res.numpy(), ref)
        hyd.stencil2d(du, hyd.STENCIL_LAPLACIAN, out=b0)  # clobber between replays
        hyd.sync()


def test_jacobi_full_size_steps(hyd, orc):
    c = workloads.JACOBI
    h, w = c["h"], c["w"]
    u = hyd.splitmix_f64(h * w, c["seed"])
    u.shape = (h, w)
    host = u.numpy()
    res, _, _ = hyd.iterate_stencil2d(u, 3, use_graph=True)
    assert bits_equal(res.numpy(), orc.iterate_stencil2d(host, 3))


def test_jacobi_200_steps_graph_equals_eager(hyd):
    c = workloads.JACOBI
    h, w = 2048, 2048
    u = hyd.splitmix_f64(h * w, c["seed"])
    u.shape = (h, w)
    g, _, _ = hyd.iterate_stencil2d(u, c["iters"], use_graph=True)
    e, _, _ = hyd.iterate_stencil2d(u, c["iters"], use_graph=False)
    assert bits_equal(g.numpy(), e.numpy())


# ---- scans, sort, gather, COO ---------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 2047, 4095, 4096, 4097, 8193, 1_000_007, (1 << 24) + 5])
@pytest.mark.parametrize("inclusive", [False, True])
@pytest.mark.parametrize("reverse", [False, True])
def test_scan(hyd, orc, n, inclusive, reverse):
    a = orc.splitmix_i64(n, 501 + n, bits=40) - (1 << 39)
    got = hyd.scan(dev(hyd, a), 11, inclusive, reverse).numpy()
    assert bits_equal(got, orc.scan(a, 11, inclusive, reverse))


@pytest.mark.parametrize("case", golden("scan"), ids=[c["name"] for c in golden("scan")])
def test_scan_known(hyd, case):
    i = case["inputs"]
    got = hyd.scan(dev(hyd, np.array(i["a"], np.int64)), i["init"], i["inclusive"], i["reverse"]).numpy()
    assert got.tolist() == case["expected"]


@pytest.mark.parametrize("n", [1, 5, 4095, 4096, 4097, 1_000_003])
@pytest.mark.parametrize("bits", [3, 8, 9, 20, 32, 40, 63])
def test_sort_indices(hyd, orc, n, bits):
    keys = orc.splitmix_i64(n, 601 + n + bits, bits=bits)
    got = hyd.sort_indices(dev(hyd, keys)).numpy()
    ref, _ = orc.sort_indices(keys)
    assert bits_equal(got, ref)


def test_sort_known_and_negative_keys(hyd, orc):
    (c,) = golden("sort_indices")
    assert hyd.sort_indices(dev(hyd, np.array(c["inputs"]["keys"], np.int64))).numpy().tolist() == c["expected"]
    # negative keys: the reference's pass count comes from max(0, max key)
    keys = orc.splitmix_i64(10_000, 606, bits=20) - (1 << 19)
    assert bits_equal(hyd.sort_indices(dev(hyd, keys)).numpy(), orc.sort_indices(keys)[0])
    # five bytes of max key: the low 40 bits are sorted (in wider digits), the
    # negative keys' upper bits ignored as in the reference's byte passes
    keys = orc.splitmix_i64(50_000, 607, bits=34) - (1 << 33)
    assert bits_equal(hyd.sort_indices(dev(hyd, keys)).numpy(), orc.sort_indices(keys)[0])


def test_gather_segmented(hyd, orc):
    n = 100_003
    src = orc.splitmix_i64(n, 701)
    perm = orc.splitmix_i64(n, 702, bits=40) % n
    assert bits_equal(hyd.gather(dev(hyd, perm), dev(hyd, src)).numpy(), orc.gather(perm, src))
    (c,) = golden("segmented_reduce")
    i = c["inputs"]
    got = hyd.segmented_reduce(dev(hyd, np.array(i["offsets"], np.int64)),
                               dev(hyd, np.array(i["vals"], np.int64)), i["init"]).numpy()
    assert got.tolist() == c["expected"]
    offs = np.sort(orc.splitmix_i64(3001, 703, bits=17))
    vals = orc.splitmix_i64(1 << 17, 704)
    got = hyd.segmented_reduce(dev(hyd, offs), dev(hyd, vals), 5).numpy()
    assert bits_equal(got, orc.segmented_reduce(offs, vals, 5))


def test_coo_known(hyd):
    (c,) = golden("coo_sum_duplicates")
    i = c["inputs"]
    r, cc, v, m = hyd.coo_sum_duplicates(*(dev(hyd, np.array(i[k], np.int64)) for k in ("rows", "cols", "vals")),
                                         i["nnz"])
    e = c["expected"]
    assert v.numpy().tolist() == e["vals"] and m == e["nnz"]
    assert r.numpy().tolist() == e["rows"] and cc.numpy().tolist() == e["cols"]
    (c,) = golden("csr_from_sorted_coo")
    i = c["inputs"]
    rp, ci, cv = hyd.csr_from_sorted_coo(i["nrows"], i["nnz"],
                                         *(dev(hyd, np.array(i[k], np.int64)) for k in ("rows", "cols", "vals")))
    e = c["expected"]
    assert rp.numpy().tolist() == e["row_ptr"] and ci.numpy().tolist() == e["col_idx"]
    assert cv.numpy().tolist() == e["vals"]
    (c,) = golden("coo_to_csr")
    i = c["inputs"]
    rp, ci, cv, m = hyd.coo_to_csr(i["nrows"], i["ncols"],
                                   *(dev(hyd, np.array(i[k], np.int64)) for k in ("rows", "cols", "vals")))
    e = c["expected"]
    assert rp.numpy().tolist() == e["row_ptr"] and ci.numpy().tolist() == e["col_idx"]
    assert cv.numpy().tolist() == e["vals"] and m == e["nnz"]


@pytest.mark.parametrize("nnz,nrows,ncols,dupbits", [(1, 1, 1, 0), (1000, 50, 40, 4),
                                                     (200_000, 1 << 20, 1 << 20, 0),
                                                     (300_000, 1000, 3, 0),
                                                     (100_000, 1_000_003, 999_983, 0)])
def test_coo_to_csr_random(hyd, orc, nnz, nrows, ncols, dupbits):
    rows = orc.splitmix_i64(nnz, 801, bits=40) % nrows
    cols = orc.splitmix_i64(nnz, 802, bits=40) % ncols
    vals = orc.splitmix_i64(nnz, 803, bits=20)
    rp, ci, cv, m = hyd.coo_to_csr(nrows, ncols, dev(hyd, rows), dev(hyd, cols), dev(hyd, vals))
    rrp, rci, rcv, rm = orc.coo_to_csr(nrows, ncols, rows, cols, vals)
    assert m == rm
    assert bits_equal(rp.numpy(), rrp) and bits_equal(ci.numpy(), rci) and bits_equal(cv.numpy(), rcv)
    # invariants: row_ptr monotone, row_ptr[nrows] == merged count
    r = rp.numpy()
    assert np.all(np.diff(r) >= 0) and r[-1] == m


@pytest.mark.parametrize("nnz,nrows,ncols", [(50_000, 2, 2), (20_000, 1, 1), (70_001, 3, 7), (4096 * 3, 1, 2)])
def test_coo_to_csr_long_runs(hyd, orc, nnz, nrows, ncols):
    # runs of equal (row, col) much longer than a 4096-element tile: tiles
    # with no run head at all, run sums carried across tiles, wrapping int64
    # sums of values near +-2^62
    rows = orc.splitmix_i64(nnz, 821, bits=40) % nrows
    cols = orc.splitmix_i64(nnz, 822, bits=40) % ncols
    vals = orc.splitmix_i64(nnz, 823, bits=63) - (1 << 62)
    print("coo long runs nnz=%d nrows=%d ncols=%d" % (nnz, nrows, ncols), flush=True)
    rp, ci, cv, m = hyd.coo_to_csr(nrows, ncols, dev(hyd, rows), dev(hyd, cols), dev(hyd, vals))
    rrp, rci, rcv, rm = orc.coo_to_csr(nrows, ncols, rows, cols, vals)
    assert m == rm
    assert bits_equal(rp.numpy(), rrp) and bits_equal(ci.numpy(), rci) and bits_equal(cv.numpy(), rcv)


@pytest.mark.parametrize("nrows,ncols,colrange", [(1000, 10, (-5, 25)), (4, 1 << 61, (0, 1 << 61))])
def test_coo_to_csr_outside_the_box(hyd, orc, nrows, ncols, colrange):
    # Columns outside [0, ncols) pack onto other rows' keys ((1, 5) and (0, 15)
    # both give 15 when ncols = 10), and nrows * ncols >= 2^62 wraps the keys:
    # both take the pair-comparing merge, as the reference's loop does.
    nnz = 20_000
    rows = orc.splitmix_i64(nnz, 811, bits=40) % nrows
    lo, hi = colrange
    cols = lo + (orc.splitmix_i64(nnz, 812, bits=62) % (hi - lo))
    cols[:4] = [5, 15, 15, 5] if ncols == 10 else cols[:4]
    rows[:4] = [1, 0, 0, 1] if ncols == 10 else rows[:4]
    vals = orc.splitmix_i64(nnz, 813, bits=63) - (1 << 62)
    print("coo outside box nrows=%d ncols=%d" % (nrows, ncols), flush=True)
    rp, ci, cv, m = hyd.coo_to_csr(nrows, ncols, dev(hyd, rows), dev(hyd, cols), dev(hyd, vals))
    rrp, rci, rcv, rm = orc.coo_to_csr(nrows, ncols, rows, cols, vals)
    assert m == rm
    assert bits_equal(rp.numpy(), rrp) and bits_equal(ci.numpy(), rci) and bits_equal(cv.numpy(), rcv)


def test_launch_shapes_leave_results_unchanged(hyd, orc):
    # the map chain and the copy at every streaming grid the tuner may pick
    # (the flat form and 1-8 workgroups per CU) give the same bits; the tuner
    # stores one class per (kind, size) and the classes are removed again
    n = (1 << 20) + 6
    xs = hyd.splitmix_f64(n, 91, 0, symmetric=True)
    ys = hyd.splitmix_f64(n, 91, n, symmetric=True)
    ref = orc.map_chain(xs.numpy(), ys.numpy())
    out = hyd.DeviceArray((n,), np.float64)
    lib = hyd.lib()
    for pc in (hyd.SHAPE_FLAT, 1, 2, 4, 8):
        print("launch shape per_cu=%d" % pc, flush=True)
        hyd.tune_shape_set(hyd.SHAPE_MAP_CHAIN, n, pc)
        hyd.tune_shape_set(hyd.SHAPE_COPY, n, pc)
        hyd.map_chain(xs, ys, out=out)
        assert np.array_equal(out.numpy().view(np.uint64), ref.view(np.uint64))
        hyd._check(lib.hydhip_copy_f64(xs.ptr, n, out.ptr, None))
        assert np.array_equal(out.numpy(), xs.numpy())
    assert (hyd.SHAPE_COPY, 20, 8) in hyd.tune_shape_table()
    for k in (hyd.SHAPE_MAP_CHAIN, hyd.SHAPE_COPY):
        hyd.tune_shape_set(k, n, 0)
    assert not [e for e in hyd.tune_shape_table() if e[1] == 20]
