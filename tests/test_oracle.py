"""Pins the CPU oracle (oracle/hyd_oracle.c) before it is trusted as the
checker: against the reference's known answers (tests/golden), closed forms,
invariants and independent implementations (numpy/scipy/pure Python) chosen
so that a dropped term, a wrong index or a transposed operand fails.
CPU only."""
import itertools
import json
import math
import os

import numpy as np
import pytest

import workloads

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def golden(kind=None):
    with open(GOLDEN) as f:
        cases = json.load(f)
    return [c for c in cases if kind is None or c["kind"] == kind]


def ids(cases):
    return [c["name"] for c in cases]


# ---- known answers --------------------------------------------------------------

@pytest.mark.parametrize("case", golden("scan"), ids=ids(golden("scan")))
def test_scan_known(orc, case):
    i = case["inputs"]
    got = orc.scan(i["a"], i["init"], i["inclusive"], i["reverse"])
    assert got.tolist() == case["expected"]


def test_segmented_reduce_known(orc):
    (c,) = golden("segmented_reduce")
    i = c["inputs"]
    assert orc.segmented_reduce(i["offsets"], i["vals"], i["init"]).tolist() == c["expected"]


def test_sort_indices_known(orc):
    (c,) = golden("sort_indices")
    perm, passes = orc.sort_indices(c["inputs"]["keys"])
    assert perm.tolist() == c["expected"]
    assert passes == 1


def test_sort_gather_known(orc):
    (c,) = golden("sort_gather")
    perm, _ = orc.sort_indices(c["inputs"]["keys"])
    assert orc.gather(perm, c["inputs"]["vals"]).tolist() == c["expected"]


def test_coo_sum_duplicates_known(orc):
    (c,) = golden("coo_sum_duplicates")
    i = c["inputs"]
    r, cc, v, m = orc.coo_sum_duplicates(i["rows"], i["cols"], i["vals"], i["nnz"])
    e = c["expected"]
    assert v.tolist() == e["vals"] and m == e["nnz"]
    assert r.tolist() == e["rows"] and cc.tolist() == e["cols"]


def test_csr_known(orc):
    (c,) = golden("csr_from_sorted_coo")
    i = c["inputs"]
    rp, ci, cv = orc.csr_from_sorted_coo(i["nrows"], i["nnz"], i["rows"], i["cols"], i["vals"])
    e = c["expected"]
    assert rp.tolist() == e["row_ptr"] and ci.tolist() == e["col_idx"] and cv.tolist() == e["vals"]


def test_coo_to_csr_known(orc):
    (c,) = golden("coo_to_csr")
    i = c["inputs"]
    rp, ci, cv, m = orc.coo_to_csr(i["nrows"], i["ncols"], i["rows"], i["cols"], i["vals"])
    e = c["expected"]
    assert rp.tolist() == e["row_ptr"] and ci.tolist() == e["col_idx"]
    assert cv.tolist() == e["vals"] and m == e["nnz"]


@pytest.mark.parametrize("case", golden("scatter"), ids=ids(golden("scatter")))
@pytest.mark.parametrize("privatized", [True, False])
def test_scatter_known(orc, case, privatized):
    i = case["inputs"]
    got = orc.scatter_add(i["idx"], i["vals"], np.zeros(i["ndst"], np.int64), privatized)
    assert got.tolist() == case["expected"]


def test_scatter_guarded_known(orc):
    (c,) = golden("scatter_guarded")
    i = c["inputs"]
    got = orc.scatter_guarded_add(i["idx"], i["vals"], i["guard"], np.array(i["defaults"], np.int64))
    assert got.tolist() == c["expected"]
    # the same answer in fp64 ((+.) form)
    got = orc.scatter_guarded_add(i["idx"], np.array(i["vals"], np.float64), i["guard"],
                                  np.array(i["defaults"], np.float64))
    assert got.tolist() == [float(v) for v in c["expected"]]


def test_scatter_guarded_matches_filtered_scatter(orc):
    # guard g == plain scatter of the elements where g holds; defaults carry through
    rng = np.random.default_rng(21)
    n, ndst = 5000, 37
    idx = rng.integers(-3, ndst + 3, n)
    vals = rng.integers(-2**40, 2**40, n)
    guard = rng.integers(0, 3, n)  # nonzero = true
    defaults = rng.integers(-1000, 1000, ndst)
    got = orc.scatter_guarded_add(idx, vals, guard, defaults)
    keep = guard != 0
    ref = defaults.copy()
    ok = keep & (idx >= 0) & (idx < ndst)
    np.add.at(ref, idx[ok], vals[ok])
    assert np.array_equal(got, ref)
    assert np.array_equal(orc.scatter_guarded_add(idx, vals, None, defaults),
                          orc.scatter_add(idx, vals, defaults, privatized=False))


def test_scatter_guarded_f64_is_serial_order(orc):
    # (+.) is not associative: the contract is the serial index order per bin.
    # 1e16 + 1 rounds back to 1e16, so ((0 + 1e16) + 1) + -1e16 = 0 whereas the
    # order 1, 1e16, -1e16 gives 0 too but 1e16, -1e16, 1 gives 1.
    idx = [0, 0, 0, 1, 1, 1]
    vals = [1e16, 1.0, -1e16, 1e16, -1e16, 1.0]
    got = orc.scatter_guarded_add(idx, np.array(vals), None, np.zeros(2))
    assert got.tolist() == [0.0, 1.0]
    # pure-Python serial loop over random data with heavy collisions
    rng = np.random.default_rng(5)
    n, ndst = 3000, 11
    idx = rng.integers(0, ndst, n)
    vals = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 9, n)
    guard = rng.integers(0, 2, n)
    dst = rng.standard_normal(ndst)
    ref = dst.copy()
    for i in range(n):
        if guard[i]:
            ref[idx[i]] = vals[i] + ref[idx[i]]
    got = orc.scatter_guarded_add(idx, vals, guard, dst)
    assert got.view(np.uint64).tolist() == ref.view(np.uint64).tolist()


def test_scatter_chain_f64_is_phase_order(orc):
    rng = np.random.default_rng(6)
    ndst = 7
    phases = []
    for k in range(3):
        m = 400 + 100 * k
        phases.append((rng.integers(0, ndst, m), rng.standard_normal(m) * 1e8,
                       None if k == 1 else rng.integers(0, 2, m)))
    dst = rng.standard_normal(ndst)
    ref = dst.copy()
    for idx, vals, guard in phases:
        for i in range(len(idx)):
            if guard is None or guard[i]:
                ref[idx[i]] = vals[i] + ref[idx[i]]
    got = orc.scatter_chain_add_f64(phases, dst)
    assert got.view(np.uint64).tolist() == ref.view(np.uint64).tolist()


def test_weighted_histogram_closed_form(orc):
    (c,) = golden("weighted_histogram")
    i = c["inputs"]
    for serial in (False, True):
        assert orc.weighted_histogram(i["data"], i["bins"], serial).tolist() == c["expected"]


@pytest.mark.parametrize("case", golden("dot"), ids=ids(golden("dot")))
def test_dot_known(orc, case):
    i = case["inputs"]
    assert orc.dot(i["xs"], i["ys"]) == case["expected"]
    assert orc.dot_serial(i["xs"], i["ys"]) == case["expected"]


@pytest.mark.parametrize("case", golden("stencil1d"), ids=ids(golden("stencil1d")))
def test_stencil1d_known(orc, case):
    i = case["inputs"]
    assert orc.stencil1d_clamp_sum(i["a"], i["offsets"]).tolist() == case["expected"]


def test_stencil2d_known(orc):
    (c,) = golden("stencil2d_jacobi4_times4")
    g = np.array(c["inputs"]["grid"], np.float64)
    out = orc.stencil2d(g, body=2)
    assert out[tuple(c["inputs"]["at"])] * 4.0 == c["expected"]
    (c,) = golden("stencil2d_laplacian")
    out = orc.stencil2d(np.array(c["inputs"]["grid"], np.float64), body=1)
    assert out[tuple(c["inputs"]["at"])] == c["expected"]


def test_jacobi_ones_fixed_point(orc):
    (c,) = golden("jacobi_iterate")
    i = c["inputs"]
    u = np.full((i["h"], i["w"]), i["value"])
    for body in (0, 2):
        assert np.all(orc.iterate_stencil2d(u, i["iters"], body=body) == c["expected"])


# ---- generator ----------------------------------------------------------------------

def test_splitmix_matches_python(orc):
    seed = workloads.HIST["seed"]
    a = orc.splitmix_i64(64, seed, counter0=5, bits=20)
    b = [workloads.splitmix_at(seed, 5 + i) >> 44 for i in range(64)]
    assert a.tolist() == b
    u = orc.splitmix_f64(64, seed, symmetric=True)
    v = [2.0 * ((workloads.splitmix_at(seed, i) >> 11) * 2.0 ** -53) - 1.0 for i in range(64)]
    assert u.tolist() == v
    # SplitMix64 reference value (Vigna's splitmix64.c, seed 0: first output)
    assert workloads.splitmix_at(0, 0) == 0xE220A8397B1DCDAF


# ---- histogram -----------------------------------------------------------------------

@pytest.mark.parametrize("nbins", [1, 7, 64, 4096, 65536])
def test_histogram_privatized_equals_serial(orc, nbins):
    data = orc.splitmix_i64(200_003, 0x1234 + nbins, bits=31)
    data[::97] = -data[::97]  # negative routes are dropped by both
    a = orc.weighted_histogram(data, nbins)
    b = orc.weighted_histogram(data, nbins, serial=True)
    assert np.array_equal(a, b)


def test_histogram_checksum_invariant(orc):
    # sum over bins = sum of x*x (mod 2^64) when every route is in range
    data = orc.splitmix_i64(100_000, 77, bits=31)
    h = orc.weighted_histogram(data, 256)
    sq = (data.astype(np.uint64) * data.astype(np.uint64)).sum(dtype=np.uint64)
    assert h.astype(np.uint64).sum(dtype=np.uint64) == sq
    # routes by brute force with numpy (a third implementation)
    ref = np.zeros(256, np.uint64)
    np.add.at(ref, data % 256, data.astype(np.uint64) * data.astype(np.uint64))
    assert np.array_equal(h.astype(np.uint64), ref)


def test_batched_histograms(orc):
    lens, bins = workloads.batch_instances(count=24, hi_log=14)
    off = np.concatenate([[0], np.cumsum(lens)[:-1]])
    data = orc.splitmix_i64(int(lens.sum()), 5, bits=31)
    hist, hoff = orc.batched_histograms(data, off, lens, bins)
    for k in range(24):
        d = data[off[k]:off[k] + lens[k]]
        assert np.array_equal(hist[hoff[k]:hoff[k] + bins[k]], orc.weighted_histogram(d, int(bins[k]), True))


# ---- dot -------------------------------------------------------------------------------

def test_dot_ones_closed_form(orc):
    n = 1 << 20
    one = np.ones(n)
    assert orc.dot(one, one) == float(n)
    assert orc.dot(one, one, init=2.5) == float(n) + 2.5


def test_dot_against_double_double(orc):
    xs = orc.splitmix_f64(1 << 22, 11, symmetric=True)
    ys = orc.splitmix_f64(1 << 22, 11, counter0=1 << 22, symmetric=True)
    exact = orc.dot_dd(xs, ys)
    # math.fsum of exactly split products: an independent exactly-rounded sum
    p = xs * ys
    e = np.array([math.fma(a, b, -c) if hasattr(math, "fma") else 0.0 for a, b, c in
                  zip(xs[:1000], ys[:1000], p[:1000])])
    assert math.fsum(list(p[:1000]) + list(e)) == pytest.approx(orc.dot_dd(xs[:1000], ys[:1000]), rel=1e-15)
    for nt in (1, 3, 8):
        assert abs(orc.dot(xs, ys, nthreads=nt) - exact) <= 1e-12 * abs(exact)
    # a dropped term is far outside the tolerance
    assert abs(orc.dot(xs[1:], ys[1:]) - exact) > 1e-12 * abs(exact)


# ---- map chain -----------------------------------------------------------------------

def chain_py(x, y):
    t1 = 2.0 * x
    t2 = t1 + y
    t3 = abs(t2)
    t4 = math.sqrt(t3)
    t5 = t4 * x
    t6 = t5 - y
    t7 = t6 if t6 > 0.0 else 0.0
    return math.exp(t7 * 1.0e-3)


def test_map_chain_against_python(orc):
    xs = orc.splitmix_f64(20_000, 21, symmetric=True)
    ys = orc.splitmix_f64(20_000, 21, counter0=20_000, symmetric=True)
    got = orc.map_chain(xs, ys)
    exp = np.array([chain_py(a, b) for a, b in zip(xs, ys)])
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64))


def test_map_chain_special_values(orc):
    xs = np.array([0.0, -0.0, 1.0, -1.0, np.nan, 0.25, 1e308, -1e308])
    ys = np.array([-0.0, 0.0, np.nan, 2.0, 1.0, -0.5, 1e308, 1.0])
    got = orc.map_chain(xs, ys)
    for g, a, b in zip(got, xs, ys):
        e = chain_py(a, b)
        assert (math.isnan(g) and math.isnan(e)) or g == e
    # NaN through `if t6 > 0.0` selects 0.0 -> exp(0) = 1 (not fmax semantics)
    assert got[2] == 1.0 and got[4] == 1.0


# ---- stencils --------------------------------------------------------------------------

def stencil_np(u, body, alpha=0.1):
    p = np.pad(u, 1, mode="edge")
    c = p[1:-1, 1:-1]
    n, s, w, e = p[:-2, 1:-1], p[2:, 1:-1], p[1:-1, :-2], p[1:-1, 2:]
    t3 = ((n + s) + w) + e
    if body == 0:
        return c + alpha * (t3 - 4.0 * c)
    if body == 1:
        return t3 + c * -4.0
    if body == 2:
        return t3 / 4.0
    return (t3 + c) / 5.0


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (7, 1), (5, 6), (67, 130)])
@pytest.mark.parametrize("body", [0, 1, 2, 3])
def test_stencil_against_numpy(orc, shape, body):
    u = orc.splitmix_f64(shape[0] * shape[1], 31 + body).reshape(shape)
    got = orc.stencil2d(u, body)
    assert np.array_equal(got.view(np.uint64), stencil_np(u, body).view(np.uint64))


def test_stencil_transposed_differs(orc):
    # the north/south vs west/east roles are not interchangeable under clamping
    u = orc.splitmix_f64(6 * 9, 3).reshape(6, 9)
    assert not np.array_equal(orc.stencil2d(u, 0), orc.stencil2d(u.T.copy(), 0).T)


def test_iterate_against_numpy(orc):
    u = orc.splitmix_f64(40 * 33, 41).reshape(40, 33)
    ref = u.copy()
    for _ in range(7):
        ref = stencil_np(ref, 0)
    got = orc.iterate_stencil2d(u, 7)
    assert np.array_equal(got, ref)
    assert np.array_equal(orc.iterate_stencil2d(u, 0), u)
    c = np.full((9, 12), 0.37)
    assert np.array_equal(orc.iterate_stencil2d(c, 5), c)  # lap(const) == 0 exactly


# ---- scans, sort, COO ------------------------------------------------------------------

@pytest.mark.parametrize("inclusive,reverse", list(itertools.product([False, True], repeat=2)))
def test_scan_against_numpy(orc, inclusive, reverse):
    a = orc.splitmix_i64(10_001, 51, bits=40) - (1 << 39)
    got = orc.scan(a, 17, inclusive, reverse)
    b = a[::-1] if reverse else a
    cs = np.cumsum(b)
    ex = cs if inclusive else np.concatenate([[0], cs[:-1]])
    ex = ex + 17
    assert np.array_equal(got, ex[::-1] if reverse else ex)


def test_segmented_reduce_against_python(orc):
    offs = np.sort(orc.splitmix_i64(50, 61, bits=10))
    vals = orc.splitmix_i64(1024, 62, bits=20)
    got = orc.segmented_reduce(offs, vals, 3)
    ex = [3 + int(vals[offs[s]:offs[s + 1]].sum()) for s in range(49)]
    assert got.tolist() == ex


@pytest.mark.parametrize("bits", [3, 8, 9, 20, 40, 62])
def test_sort_against_stable_argsort(orc, bits):
    keys = orc.splitmix_i64(30_000, 71 + bits, bits=bits)
    perm, passes = orc.sort_indices(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    assert passes == max(1, (int(keys.max()).bit_length() + 7) // 8)


def test_coo_sum_duplicates_against_groupby(orc):
    rows = np.sort(orc.splitmix_i64(5000, 81, bits=6))
    cols = orc.splitmix_i64(5000, 82, bits=2)
    vals = orc.splitmix_i64(5000, 83, bits=30)
    r, c, v, m = orc.coo_sum_duplicates(rows, cols, vals)
    groups = [(k, sum(x[2] for x in g)) for k, g in
              itertools.groupby(zip(rows.tolist(), cols.tolist(), vals.tolist()), key=lambda t: t[:2])]
    assert m == len(groups)
    assert r[:m].tolist() == [k[0] for k, _ in groups] and c[:m].tolist() == [k[1] for k, _ in groups]
    assert v[:m].tolist() == [s for _, s in groups]
    assert not r[m:].any() and not c[m:].any() and not v[m:].any()


def test_coo_to_csr_against_scipy(orc):
    sp = pytest.importorskip("scipy.sparse")
    nrows, ncols, nnz = 300, 200, 20_000
    rows = orc.splitmix_i64(nnz, 91, bits=40) % nrows
    cols = orc.splitmix_i64(nnz, 92, bits=40) % ncols
    vals = orc.splitmix_i64(nnz, 93, bits=20)
    rp, ci, cv, m = orc.coo_to_csr(nrows, ncols, rows, cols, vals)
    ref = sp.coo_matrix((vals, (rows, cols)), shape=(nrows, ncols)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(rp, ref.indptr.astype(np.int64))
    assert np.array_equal(ci[:m], ref.indices) and np.array_equal(cv[:m], ref.data)
    assert m == ref.nnz and not ci[m:].any() and not cv[m:].any()
