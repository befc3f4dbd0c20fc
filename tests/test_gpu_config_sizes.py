"""GPU parity at the benchmark configurations' own sizes.

* the served mix drawn from its full ranges (stencil side up to 4096 with
  10-100 steps, 64-bin histograms up to 2^24, dots up to 2^26, COO->CSR
  builds of 1e5-1e7 nonzeros over 2^20 x 2^20), each class's lower and upper
  end included: every result equals its alone run bit for bit and matches the
  oracle (integers and stencils bit-exact, dots within 1e-12 relative);
* COO->CSR of 1e7 nonzeros over 2^20 x 2^20 and sort_indices of 1e7 40-bit
  keys (the COO build's packed keys), bit-exact;
* the 8192 x 8192 Jacobi heat grid, 200 steps in one hipGraph, bit-exact
  against the oracle's 200 serial steps.
"""
import numpy as np
import pytest

import servemix
import workloads

pytestmark = pytest.mark.gpu

REL_TOL_REDUCTION = 1e-12


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float64:
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return a.shape == b.shape and np.array_equal(a, b)


def same(a, b):
    if isinstance(a, tuple):
        return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    if isinstance(a, (int, np.integer)):
        return a == b
    return bits_equal(np.asarray(a), np.asarray(b))


def config_mix():
    """56 requests from the served mix's full ranges plus both ends of every
    class."""
    reqs = workloads.serve_mix(56, seed=workloads.SERVE["seed"] + 300)
    ends = [dict(kind="stencil", side=4096, iters=100), dict(kind="stencil", side=256, iters=10),
            dict(kind="hist", n=1 << 24, bins=64), dict(kind="hist", n=1 << 12, bins=64),
            dict(kind="dot", n=1 << 26), dict(kind="dot", n=1 << 14),
            dict(kind="coo", nnz=10_000_000, nrows=1 << 20, ncols=1 << 20),
            dict(kind="coo", nnz=100_000, nrows=1 << 20, ncols=1 << 20)]
    # interleave the ends so that they run beside other classes in the mix
    for k, r in enumerate(ends):
        reqs.insert(7 * k + 3, r)
    return reqs


def oracle_result(orc, r, host):
    k = r["kind"]
    if k == "stencil":
        u = host["grid"][: r["side"] ** 2].reshape(r["side"], r["side"])
        return orc.iterate_stencil2d(u, r["iters"], 0, 0.1).ravel()
    if k == "hist":
        return orc.weighted_histogram(host["data"][: r["n"]], r["bins"])
    if k == "dot":
        return np.array([orc.dot(host["xs"][: r["n"]], host["ys"][: r["n"]])])
    n = r["nnz"]
    return orc.coo_to_csr(r["nrows"], r["ncols"], host["rows"][:n], host["cols"][:n], host["vals"][:n])


def test_served_mix_config_ranges(hyd, orc):
    reqs = config_mix()
    kinds = {k: [r for r in reqs if r["kind"] == k] for k in ("stencil", "hist", "dot", "coo")}
    assert max(r["side"] for r in kinds["stencil"]) == 4096
    assert max(r["n"] for r in kinds["hist"]) == 1 << 24 and max(r["n"] for r in kinds["dot"]) == 1 << 26
    assert max(r["nnz"] for r in kinds["coo"]) == 10_000_000 and min(r["nnz"] for r in kinds["coo"]) == 100_000
    pools = servemix.Pools(reqs)
    host = dict(grid=pools.grid.numpy(), data=pools.data.numpy(), xs=pools.xs.numpy(), ys=pools.ys.numpy(),
                rows=pools.rows.numpy(), cols=pools.cols.numpy(), vals=pools.vals.numpy())
    budget = 6 << 30
    alone = []
    srv = hyd.Server(budget, workers=1, flags=hyd.SERVE_NO_BATCH)
    try:
        for k, r in enumerate(reqs):
            print("alone %d/%d %s" % (k + 1, len(reqs), r), flush=True)
            rid = srv.submit(pools.request(r))
            res = srv.wait(rid, 120_000.0)
            assert res.status == hyd.HYD_OK, r
            alone.append(servemix.fetch(r, res))
            srv.release(rid)
    finally:
        srv.close()
    print("mix of %d requests" % len(reqs), flush=True)
    mix = hyd.Server(budget, workers=4)
    try:
        ids = [mix.submit(pools.request(r)) for r in reqs]  # all at once: overlap + batching
        assert mix.wait_all(240_000)
        st = mix.stats()
        print("mix stats", st, flush=True)
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["batched_jobs"] >= 1
        for r, rid, a in zip(reqs, ids, alone):
            res = mix.wait(rid)
            assert res.status == hyd.HYD_OK
            assert same(servemix.fetch(r, res), a), r  # bit-identical alone and in the mix
            mix.release(rid)
        assert mix.stats()["in_use"] == 0
    finally:
        mix.close()
        pools.free()
    for k, (r, a) in enumerate(zip(reqs, alone)):
        print("oracle %d/%d %s" % (k + 1, len(reqs), r), flush=True)
        ref = oracle_result(orc, r, host)
        if r["kind"] == "dot":
            assert abs(a[0] - ref[0]) <= REL_TOL_REDUCTION * abs(ref[0]), r
        else:
            assert same(a, ref), r


def test_coo_to_csr_1e7(hyd, orc):
    nnz, nrows = 10_000_000, 1 << 20
    seed = workloads.SERVE["seed"] + 13
    rows = hyd.splitmix_i64(nnz, seed, bits=20)
    cols = hyd.splitmix_i64(nnz, seed + 1, bits=20)
    vals = hyd.splitmix_i64(nnz, seed + 2, bits=20)
    print("coo_to_csr nnz=%d" % nnz, flush=True)
    rp, ci, cv, m = hyd.coo_to_csr(nrows, nrows, rows, cols, vals)
    hr, hc, hv = rows.numpy(), cols.numpy(), vals.numpy()
    rrp, rci, rcv, rm = orc.coo_to_csr(nrows, nrows, hr, hc, hv)
    assert m == rm
    assert bits_equal(rp.numpy(), rrp) and bits_equal(ci.numpy(), rci) and bits_equal(cv.numpy(), rcv)
    r = rp.numpy()
    # invariants: monotone row_ptr ending at the merged count; the merged
    # values keep the total (no value lost or counted twice)
    assert np.all(np.diff(r) >= 0) and r[-1] == m and r[0] == 0
    assert int(cv.numpy()[:m].sum()) == int(hv.sum())
    # 2^40 packed keys over 1e7 draws: duplicates exist and were merged
    assert m < nnz
    for a in (rows, cols, vals, rp, ci, cv):
        a.free()


def test_sort_indices_1e7_40bit(hyd, orc):
    n = 10_000_000
    seed = workloads.SERVE["seed"] + 16
    keys = hyd.splitmix_i64(n, seed, bits=40)
    print("sort_indices n=%d 40-bit keys" % n, flush=True)
    perm = hyd.sort_indices(keys).numpy()
    hk = keys.numpy()
    ref, _ = orc.sort_indices(hk)
    assert bits_equal(perm, ref)
    # a stable permutation: sorted keys, ties in index order
    sk = hk[perm]
    assert np.all(np.diff(sk) >= 0)
    tie = np.diff(sk) == 0
    assert np.all(np.diff(perm)[tie] > 0)
    assert np.array_equal(np.sort(perm), np.arange(n))
    keys.free()


def test_jacobi_config_200_steps_graph(hyd, orc):
    c = workloads.JACOBI
    h, w, iters = c["h"], c["w"], c["iters"]
    u = hyd.splitmix_f64(h * w, c["seed"])
    u.shape = (h, w)
    print("jacobi %dx%d x %d steps, one graph" % (h, w, iters), flush=True)
    res, b0, b1 = hyd.iterate_stencil2d(u, iters, alpha=c["alpha"], use_graph=True)
    got = res.numpy()
    host = u.numpy()
    for a in (u, b0, b1):
        a.free()
    ref = orc.iterate_stencil2d(host, iters, 0, c["alpha"])
    assert bits_equal(got, ref)
