"""CPU tests of the serving scheduler (include/hyd_serve.h) with the fake
device: admission by HBM budget, rejection, deadlines, injected failures,
continuous batching.  No GPU calls are made in this mode."""
import time

import pytest

import hydhip

FAKE_IN = 1 << 20  # any non-null address: the fake device never dereferences it


def stencil(side, iters, **kw):
    return hydhip.request(hydhip.REQ_STENCIL, h=side, w=side, iters=iters, in0=FAKE_IN, **kw)


def hist(n, bins=64, **kw):
    return hydhip.request(hydhip.REQ_HIST, n=n, bins=bins, in0=FAKE_IN, **kw)


def dot(n, **kw):
    return hydhip.request(hydhip.REQ_DOT, n=n, in0=FAKE_IN, in1=FAKE_IN, **kw)


def coo(nnz, nrows=1 << 20, **kw):
    return hydhip.request(hydhip.REQ_COO, n=nnz, nrows=nrows, ncols=nrows, in0=FAKE_IN, in1=FAKE_IN,
                          in2=FAKE_IN, **kw)


@pytest.fixture
def fake():
    servers = []

    def make(budget, workers=2, flags=0):
        s = hydhip.Server(budget, workers, flags | hydhip.SERVE_FAKE)
        servers.append(s)
        return s

    yield make
    for s in servers:
        s.close()


def test_request_bytes():
    assert hydhip.request_bytes(hist(1000, 64)) == 512
    assert hydhip.request_bytes(stencil(256, 10)) == 2 * 256 * 256 * 8
    assert hydhip.request_bytes(stencil(256, 1)) == 256 * 256 * 8
    assert hydhip.request_bytes(dot(10)) == 256
    assert hydhip.request_bytes(coo(1000, 100)) > 3 * 1000 * 8
    bad = hydhip.request(hydhip.REQ_HIST, n=5, bins=0, in0=FAKE_IN)
    assert hydhip.request_bytes(bad) == -1


def test_mixed_requests_complete(fake):
    s = fake(1 << 30, workers=3)
    ids = [s.submit(r) for r in (stencil(256, 10), hist(5000), dot(7000), coo(10000), hist(3), dot(0))]
    assert s.wait_all(10_000)
    for rid in ids:
        res = s.wait(rid)
        assert res.done == 1 and res.status == hydhip.HYD_OK and res.out0
        assert res.submit_ms <= res.admit_ms <= res.start_ms <= res.done_ms
        s.release(rid)
    st = s.stats()
    assert st["completed"] == 6 and st["in_use"] == 0 and st["failed"] == 0


def test_rejects_requests_larger_than_budget(fake):
    s = fake(1 << 20)
    with pytest.raises(hydhip.HydError) as e:
        s.submit(stencil(1024, 10))  # 16 MiB > 1 MiB
    assert e.value.code == hydhip.HYD_E_REJECTED
    assert s.stats()["rejected"] == 1
    with pytest.raises(hydhip.HydError) as e:
        s.submit(hydhip.request(hydhip.REQ_DOT, n=10))  # null inputs
    assert e.value.code == hydhip.HYD_E_INVALID


def test_admission_respects_budget(fake):
    budget = 1 << 20  # 1 MiB: two 512 KiB stencils at a time
    s = fake(budget, workers=4)
    ids = [s.submit(stencil(128, 50)) for _ in range(12)]  # 2 x 128 KiB each while running
    for rid in ids:
        res = s.wait(rid, 10_000)
        assert res.status == hydhip.HYD_OK
        s.release(rid)
    st = s.stats()
    assert st["completed"] == 12 and st["peak_in_use"] <= budget and st["in_use"] == 0


def test_outputs_hold_memory_until_release(fake):
    s = fake(4 << 20, workers=2)
    a = s.submit(stencil(512, 5))  # 2 MiB while running, 2 MiB kept
    res = s.wait(a)
    assert res.bytes == 512 * 512 * 8
    assert s.stats()["in_use"] == 512 * 512 * 8
    s.release(a)
    assert s.stats()["in_use"] == 0


def test_deadline_expires_while_blocked(fake):
    s = fake(4 << 20, workers=2)
    big = s.submit(stencil(512, 5))
    s.wait(big)  # keeps 2 MiB until released
    blocked = s.submit(stencil(512, 50, deadline_ms=20.0))  # needs 4 MiB: cannot be admitted
    time.sleep(0.1)
    res = s.wait(blocked, 5_000)
    assert res.status == hydhip.HYD_E_TIMEOUT and res.out0 is None
    assert s.stats()["timed_out"] == 1
    # the server keeps working afterwards
    s.release(big)
    ok = s.submit(stencil(512, 50))
    assert s.wait(ok, 5_000).status == hydhip.HYD_OK


def test_injected_failure_fails_one_job_only(fake):
    s = fake(1 << 30, workers=1, flags=hydhip.SERVE_NO_BATCH)
    s.inject_failures(1)
    ids = [s.submit(dot(1000)) for _ in range(4)]
    res = [s.wait(r) for r in ids]
    statuses = [r.status for r in res]
    assert statuses.count(hydhip.HYD_E_HIP) == 1 and statuses.count(hydhip.HYD_OK) == 3
    failed = [r for r in res if r.status != hydhip.HYD_OK][0]
    assert failed.out0 is None and failed.bytes == 0
    for r in ids:
        s.release(r)
    assert s.stats()["in_use"] == 0


def test_continuous_batching_groups_waiting_requests(fake):
    s = fake(1 << 30, workers=1)
    # occupies the only worker for ~27 ms of modelled device time (16 B per
    # cell and step at 4 TB/s), far longer than submitting the 30 below
    first = s.submit(stencil(4096, 400))
    ids = [s.submit(hist(1 << 14)) for _ in range(20)] + [s.submit(dot(1 << 14)) for _ in range(10)]
    assert s.wait_all(30_000)
    res = [s.wait(r) for r in ids]
    hist_batches = {r.batch_id for r in res[:20]}
    dot_batches = {r.batch_id for r in res[20:]}
    assert len(hist_batches) == 1 and len(dot_batches) == 1  # arrived while busy: one launch each
    st = s.stats()
    assert st["batched_jobs"] >= 2 and st["batched_requests"] >= 30
    for r in [first] + ids:
        s.release(r)


def test_no_batch_flag(fake):
    s = fake(1 << 30, workers=1, flags=hydhip.SERVE_NO_BATCH)
    s.submit(stencil(1024, 100))
    ids = [s.submit(hist(1 << 10)) for _ in range(5)]
    assert s.wait_all(30_000)
    assert len({s.wait(r).batch_id for r in ids}) == 5
    assert s.stats()["batched_jobs"] == 0


@pytest.mark.parametrize("req,nbufs", [(lambda: stencil(64, 5), 2), (lambda: stencil(64, 1), 1),
                                       (lambda: hist(100), 1), (lambda: dot(100), 1),
                                       (lambda: coo(1000, 100), 3)])
def test_plain_path_frees_on_allocation_failure(req, nbufs):
    # each allocation of the plain per-call path fails in turn: whatever was
    # allocated before it is freed and the failure is reported
    import ctypes

    r = req()
    live = ctypes.c_int64()
    L = hydhip.lib()
    assert L.hydhip_test_plain_alloc(ctypes.byref(r), -1, ctypes.byref(live)) == hydhip.HYD_OK
    assert live.value == nbufs
    for k in range(nbufs):
        rc = L.hydhip_test_plain_alloc(ctypes.byref(r), k, ctypes.byref(live))
        assert rc == hydhip.HYD_E_OOM, k
        assert live.value == 0, k
        assert b"injected" in L.hydhip_last_error()
