"""GPU tests of the serving layer: each request of a mixed stream gives the
same result alone and in the mix (bit-exact, dots included: the batched dot
sums each request in a fixed order), and matches the oracle (integers and
stencils bit-exact, dots within 1e-12 relative)."""
import numpy as np
import pytest

import servemix
import workloads

pytestmark = pytest.mark.gpu

REL_TOL_REDUCTION = 1e-12


def small_mix(count, seed):
    return workloads.serve_mix(count, seed=seed, max_side_log=9, max_hist_log=16, max_dot_log=16,
                               nnz_lo=1e3, nnz_hi=5e4)


def oracle_result(orc, r, pools):
    k = r["kind"]
    if k == "stencil":
        u = pools.grid.numpy()[: r["side"] ** 2].reshape(r["side"], r["side"])
        return orc.iterate_stencil2d(u, r["iters"], 0, 0.1).ravel()
    if k == "hist":
        return orc.weighted_histogram(pools.data.numpy()[: r["n"]], r["bins"])
    if k == "dot":
        return np.array([orc.dot(pools.xs.numpy()[: r["n"]], pools.ys.numpy()[: r["n"]])])
    n = r["nnz"]
    rp, ci, cv, m = orc.coo_to_csr(r["nrows"], r["ncols"], pools.rows.numpy()[:n], pools.cols.numpy()[:n],
                                   pools.vals.numpy()[:n])
    return rp, ci, cv, m


def same(a, b):
    if isinstance(a, tuple):
        return all(same(x, y) for x, y in zip(a, b))
    if isinstance(a, (int, np.integer)):
        return a == b
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def check_oracle(r, got, ref):
    if r["kind"] == "dot":
        assert abs(got[0] - ref[0]) <= REL_TOL_REDUCTION * abs(ref[0])
    else:
        assert same(got, ref), r


def test_each_kind_alone_matches_oracle(hyd, orc):
    reqs = [dict(kind="stencil", side=200, iters=13), dict(kind="hist", n=50_001, bins=64),
            dict(kind="dot", n=70_001), dict(kind="coo", nnz=30_000, nrows=1 << 20, ncols=1 << 20),
            dict(kind="stencil", side=256, iters=0), dict(kind="stencil", side=128, iters=1),
            dict(kind="hist", n=0, bins=64), dict(kind="dot", n=0)]
    pools = servemix.Pools(reqs)
    srv = hyd.Server(1 << 30, workers=2)
    try:
        for r in reqs:
            print("serve alone", r, flush=True)
            rid = srv.submit(pools.request(r))
            res = srv.wait(rid)
            assert res.status == hyd.HYD_OK
            check_oracle(r, servemix.fetch(r, res), oracle_result(orc, r, pools))
            srv.release(rid)
    finally:
        srv.close()
        pools.free()


def test_mix_equals_alone(hyd, orc):
    reqs = small_mix(80, seed=workloads.SERVE["seed"] + 100)
    pools = servemix.Pools(reqs)
    alone = []
    srv = hyd.Server(2 << 30, workers=1, flags=hyd.SERVE_NO_BATCH)
    try:
        for r in reqs:
            rid = srv.submit(pools.request(r))
            res = srv.wait(rid)
            assert res.status == hyd.HYD_OK
            alone.append(servemix.fetch(r, res))
            srv.release(rid)
    finally:
        srv.close()
    mix = hyd.Server(2 << 30, workers=4)
    try:
        ids = [mix.submit(pools.request(r)) for r in reqs]  # all at once: overlap + batching
        assert mix.wait_all(120_000)
        st = mix.stats()
        print("mix stats", st, flush=True)
        assert st["completed"] == len(reqs) and st["batched_jobs"] >= 1
        for r, rid, a in zip(reqs, ids, alone):
            res = mix.wait(rid)
            assert res.status == hyd.HYD_OK
            got = servemix.fetch(r, res)
            assert same(got, a), r  # bit-identical alone and in the mix
            mix.release(rid)
        assert mix.stats()["in_use"] == 0
    finally:
        mix.close()
    for r, a in zip(reqs, alone):
        check_oracle(r, a, oracle_result(orc, r, pools))
    pools.free()


def test_plain_path_matches_oracle(hyd, orc):
    reqs = small_mix(16, seed=workloads.SERVE["seed"] + 200)
    pools = servemix.Pools(reqs)
    for r in reqs:
        res = hyd.run_plain(pools.request(r))
        check_oracle(r, servemix.fetch(r, res), oracle_result(orc, r, pools))
        hyd.free_plain(res)
    pools.free()


def test_budget_limits_concurrency_on_gpu(hyd):
    reqs = [dict(kind="stencil", side=512, iters=20) for _ in range(8)]
    pools = servemix.Pools(reqs)
    budget = 5 << 20  # two running stencils (4 MiB each with the partner) never fit together
    srv = hyd.Server(budget, workers=4)
    try:
        ids = [srv.submit(pools.request(r)) for r in reqs]
        first = None
        for rid in ids:
            res = srv.wait(rid)
            assert res.status == hyd.HYD_OK
            got = servemix.fetch(reqs[0], res)
            first = got if first is None else first
            assert same(got, first)
            srv.release(rid)
        st = srv.stats()
        assert st["peak_in_use"] <= budget and st["completed"] == 8
    finally:
        srv.close()
        pools.free()
