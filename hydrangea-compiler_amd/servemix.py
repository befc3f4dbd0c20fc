"""Device-resident input pools and hyd_request_t construction for the served
mix (shared by bench.py and the serving tests).  Every request reads a
prefix of one pool, so inputs are generated once and stay in HBM."""
import numpy as np

import hydhip
import workloads


class Pools:
    def __init__(self, reqs, seed=workloads.SERVE["seed"]):
        side = max([r["side"] for r in reqs if r["kind"] == "stencil"] or [1])
        nh = max([r["n"] for r in reqs if r["kind"] == "hist"] or [1])
        nd = max([r["n"] for r in reqs if r["kind"] == "dot"] or [1])
        nc = max([r["nnz"] for r in reqs if r["kind"] == "coo"] or [1])
        self.grid = hydhip.splitmix_f64(side * side, seed + 10)
        self.data = hydhip.splitmix_i64(nh, seed + 11, bits=20)
        self.xs = hydhip.splitmix_f64(nd, seed + 12, symmetric=True)
        self.ys = hydhip.splitmix_f64(nd, seed + 12, counter0=nd, symmetric=True)
        self.rows = hydhip.splitmix_i64(nc, seed + 13, bits=20)
        self.cols = hydhip.splitmix_i64(nc, seed + 14, bits=20)
        self.vals = hydhip.splitmix_i64(nc, seed + 15, bits=20)
        hydhip.device_sync()

    def request(self, r, deadline_ms=0.0):
        k = r["kind"]
        if k == "stencil":
            return hydhip.request(hydhip.REQ_STENCIL, h=r["side"], w=r["side"], iters=r["iters"],
                                  in0=self.grid.ptr, deadline_ms=deadline_ms)
        if k == "hist":
            return hydhip.request(hydhip.REQ_HIST, n=r["n"], bins=r["bins"], in0=self.data.ptr,
                                  deadline_ms=deadline_ms)
        if k == "dot":
            return hydhip.request(hydhip.REQ_DOT, n=r["n"], in0=self.xs.ptr, in1=self.ys.ptr,
                                  deadline_ms=deadline_ms)
        return hydhip.request(hydhip.REQ_COO, n=r["nnz"], nrows=r["nrows"], ncols=r["ncols"],
                              in0=self.rows.ptr, in1=self.cols.ptr, in2=self.vals.ptr,
                              deadline_ms=deadline_ms)

    def free(self):
        for a in (self.grid, self.data, self.xs, self.ys, self.rows, self.cols, self.vals):
            a.free()


def fetch(r, res):
    """Host copy of a finished request's outputs (numpy)."""
    k = r["kind"]
    if k == "stencil":
        return hydhip.copy_out(res.out0, r["side"] ** 2, np.float64)
    if k == "hist":
        return hydhip.copy_out(res.out0, r["bins"], np.int64)
    if k == "dot":
        return hydhip.copy_out(res.out0, 1, np.float64)
    return (hydhip.copy_out(res.out0, r["nrows"] + 1, np.int64),
            hydhip.copy_out(res.out1, r["nnz"], np.int64),
            hydhip.copy_out(res.out2, r["nnz"], np.int64), res.count)
