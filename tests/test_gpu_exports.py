"""GPU tests of the drop-in boundary: hyd_export_* called through the C ABI
with hyd_array_t arguments exactly as a host linked against the reference's
exported header would (demo/gray_scott/main.cpp:85-119), checked against the
oracle."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture
def lib(hyd):
    l = hyd.lib()
    l.hyd_hip_set_host_sync(1)
    return l


def test_export_weighted_histogram(hyd, orc, lib):
    data = orc.splitmix_i64(1 << 18, workloads.HIST["seed"], bits=20)
    arr = hyd.host_array(data)
    out = lib.hyd_export_weighted_histogram(arr)
    assert out.contents.shape.ndims == 1 and out.contents.shape.elems[0] == 64
    assert np.array_equal(hyd.read_array(out, np.int64), orc.weighted_histogram(data, 64))
    out2 = lib.hyd_export_weighted_histogram_bins(arr, 4096)
    assert np.array_equal(hyd.read_array(out2, np.int64), orc.weighted_histogram(data, 4096))
    lib.hyd_array_free(out)
    lib.hyd_array_free(out2)
    lib.hyd_array_free(arr)


@pytest.mark.parametrize("n", [(1 << 20) + 3, (1 << 26) + 12_345])
def test_export_large_host_arrays(hyd, orc, lib, n):
    # inputs of 8 MiB and more are uploaded, and results downloaded, through
    # the pinned 256 MiB halves (the second size spans three of them, the
    # last one partial): the drop-in results stay bit-exact
    print("export large host arrays n=%d" % n, flush=True)
    xs = orc.splitmix_f64(n, 5, symmetric=True)
    ys = orc.splitmix_f64(n, 5, counter0=n, symmetric=True)
    ax, ay = hyd.host_array(xs), hyd.host_array(ys)
    m = lib.hyd_export_map_chain(ax, ay)
    got = hyd.read_array(m, np.float64)
    assert np.array_equal(got.view(np.uint64), orc.map_chain(xs, ys).view(np.uint64))
    d = lib.hyd_export_dot(ax, ay)
    ref = orc.dot(xs, ys)
    assert abs(d - ref) <= 1e-12 * abs(ref)
    for x in (m, ax, ay):
        lib.hyd_array_free(x)


def test_export_input_rewritten_between_calls(hyd, orc, lib):
    # host copy of an input is authoritative: rewriting arr->data is seen
    data = np.arange(1000, dtype=np.int64)
    arr = hyd.host_array(data)
    a = lib.hyd_export_weighted_histogram(arr)
    np.ctypeslib.as_array(arr.contents.data, shape=(8000,))[:] = np.frombuffer((data * 3).tobytes(), np.uint8)
    b = lib.hyd_export_weighted_histogram(arr)
    assert np.array_equal(hyd.read_array(a, np.int64), orc.weighted_histogram(data, 64))
    assert np.array_equal(hyd.read_array(b, np.int64), orc.weighted_histogram(data * 3, 64))
    for x in (a, b, arr):
        lib.hyd_array_free(x)


def test_export_dot_and_map_chain(hyd, orc, lib):
    n = 100_001
    xs = orc.splitmix_f64(n, 3, symmetric=True)
    ys = orc.splitmix_f64(n, 3, counter0=n, symmetric=True)
    ax, ay = hyd.host_array(xs), hyd.host_array(ys)
    d = lib.hyd_export_dot(ax, ay)
    ref = orc.dot(xs, ys)
    assert abs(d - ref) <= 1e-12 * abs(ref)
    m = lib.hyd_export_map_chain(ax, ay)
    got = hyd.read_array(m, np.float64)
    assert np.array_equal(got.view(np.uint64), orc.map_chain(xs, ys).view(np.uint64))
    for x in (m, ax, ay):
        lib.hyd_array_free(x)


def test_export_stencils_and_iterate(hyd, orc, lib):
    u = orc.splitmix_f64(96 * 130, 5).reshape(96, 130)
    au = hyd.host_array(u)
    for n in (0, 1, 2, 7):
        r = lib.hyd_export_jacobi_heat(au, n)
        assert np.array_equal(hyd.read_array(r, np.float64), orc.iterate_stencil2d(u, n, 0, 0.1))
        lib.hyd_array_free(r)
    s = lib.hyd_export_heat_step(au)
    assert np.array_equal(hyd.read_array(s, np.float64), orc.stencil2d(u, 0, 0.1))
    lp = lib.hyd_export_laplacian(au)
    assert np.array_equal(hyd.read_array(lp, np.float64), orc.stencil2d(u, 1))
    for x in (s, lp, au):
        lib.hyd_array_free(x)


def test_export_coo_to_csr_record(hyd, orc, lib):
    nrows, ncols, nnz = 1000, 700, 50_000
    rows = orc.splitmix_i64(nnz, 11, bits=40) % nrows
    cols = orc.splitmix_i64(nnz, 12, bits=40) % ncols
    vals = orc.splitmix_i64(nnz, 13, bits=20)
    ar, ac, av = hyd.host_array(rows), hyd.host_array(cols), hyd.host_array(vals)
    rec = lib.hyd_export_coo_to_csr(nrows, ncols, nnz, ar, ac, av)
    rp, ci, cv, m = orc.coo_to_csr(nrows, ncols, rows, cols, vals)
    assert rec.nnz == m and rec.nrows == nrows and rec.ncols == ncols
    assert np.array_equal(hyd.read_array(rec.row_ptr, np.int64), rp)
    assert np.array_equal(hyd.read_array(rec.col_idx, np.int64), ci)
    assert np.array_equal(hyd.read_array(rec.vals, np.int64), cv)
    for x in (rec.row_ptr, rec.col_idx, rec.vals, ar, ac, av):
        lib.hyd_array_free(x)


def test_sort_indices_runtime_function(hyd, orc, lib):
    keys = orc.splitmix_i64(70_001, 17, bits=33)
    ak = hyd.host_array(keys)
    p = lib.hyd_sort_indices(ak)
    assert np.array_equal(hyd.read_array(p, np.int64), orc.sort_indices(keys)[0])
    lib.hyd_array_free(p)
    lib.hyd_array_free(ak)


def test_device_resident_chain_without_host_sync(hyd, orc, lib):
    # results stay in HBM between calls; the host copy is refreshed on demand
    u = orc.splitmix_f64(64 * 64, 19).reshape(64, 64)
    au = hyd.host_array(u)
    lib.hyd_hip_set_host_sync(0)
    try:
        a = lib.hyd_export_heat_step(au)
        assert lib.hyd_array_is_device_resident(a) == 1
        b = lib.hyd_export_heat_step(a)  # reads a's HBM copy in place
        hyd._check(lib.hyd_array_sync_host(b))
        assert np.array_equal(hyd.read_array(b, np.float64), orc.iterate_stencil2d(u, 2))
        c = lib.hyd_array_alloc_copy(b)
        hyd._check(lib.hyd_array_sync_host(c))
        assert np.array_equal(hyd.read_array(c, np.float64), hyd.read_array(b, np.float64))
        for x in (a, b, c):
            lib.hyd_array_free(x)
    finally:
        lib.hyd_hip_set_host_sync(1)
        lib.hyd_array_free(au)
