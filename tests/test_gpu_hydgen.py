"""GPU parity of the kernels hydgen generates for the example programs
(lib/libhydgen_programs.so), called through their C-ABI launchers with
hyd_array_t arguments, as the printed C calls them.

Oracle: the serial CFG interpreter (oracle/cfg_interp.py, pinned on the CPU)
and, where the C oracle has the same program, the C oracle at the benchmark
sizes.  Bars: integer and index results bit-exact; elementwise and stencil
fp64 bit-exact; fp64 reductions within 1e-12 relative; the order-exact fp64
scatter bit-exact to the serial loop.
"""
import numpy as np
import pytest

import cfg_harness as H
import workloads

pytestmark = pytest.mark.gpu

REL_TOL_REDUCTION = 1e-12


@pytest.fixture(scope="module")
def prog(hyd):
    import hydgen_lib

    return hydgen_lib.programs()


class Arrays:
    """hyd_array_t* arguments made from numpy arrays (host copies, uploaded
    by the first launcher that reads them), read back after the launchers."""

    def __init__(self, hyd):
        self.hyd = hyd
        self.L = hyd.lib()
        self.made = []

    def host(self, a):
        p = self.hyd.host_array(np.ascontiguousarray(a))
        self.made.append(p)
        return p

    def device(self, shape):
        p = self.L.hyd_array_alloc_device8(self.hyd.hyd_tuple(*shape))
        self.made.append(p)
        return p

    def read(self, p, dtype):
        assert self.L.hyd_array_sync_host(p) == 0
        return self.hyd.read_array(p, dtype)

    def free(self):
        for p in self.made:
            self.L.hyd_array_free(p)
        self.made = []


@pytest.fixture
def arrays(hyd):
    a = Arrays(hyd)
    yield a
    a.free()


def call(hyd, prog, s, env):
    """The site's launcher with env's values in its argument order."""
    args = []
    for v, t in s.get("hsArgs"):
        x = env[v]
        if t.name == "CTTuple":
            x = hyd.hyd_tuple(*x)
        args.append(x)
    return getattr(prog, s.get("hsName"))(*args)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


SIZES = [0, 1, 1023, 1024, 1025, 100_003]


@pytest.mark.parametrize("n", SIZES)
def test_map_chain(hyd, orc, prog, arrays, n):
    xs = orc.splitmix_f64(n, 901 + n, symmetric=True)
    ys = orc.splitmix_f64(n, 902 + n, symmetric=True)
    out = arrays.device((n,))
    call(hyd, prog, H.site("map_chain"), dict(n=n, out=out, xs=arrays.host(xs), ys=arrays.host(ys)))
    assert hyd.lib().hyd_array_is_device_resident(out) == 1  # left in HBM: no copy back per loop
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(orc.map_chain(xs, ys)))


def test_map_chain_full_size(hyd, orc, prog, arrays):
    c = workloads.DOTMAP
    n = c["n"]
    xs = orc.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = orc.splitmix_f64(n, c["seed"], n, symmetric=True)
    out = arrays.device((n,))
    print("generated map chain n=%d" % n, flush=True)
    call(hyd, prog, H.site("map_chain"), dict(n=n, out=out, xs=arrays.host(xs), ys=arrays.host(ys)))
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(orc.map_chain(xs, ys)))


@pytest.mark.parametrize("n", [1, 1025, 30_001])
def test_map_chain2(hyd, prog, arrays, n):
    xs, ks = H.map_chain2_inputs(n, 77 + n)
    of, oi = arrays.device((n,)), arrays.device((n,))
    env = dict(n=n, xs=arrays.host(xs), ks=arrays.host(ks), of=of, oi=oi, scale=7.25, m=11)
    call(hyd, prog, H.site("map_chain2"), env)
    u, g2 = H.map_chain2_numpy(xs, ks, 7.25, 11)
    assert np.array_equal(arrays.read(oi, np.int64), g2)
    assert np.array_equal(bits(arrays.read(of, np.float64)), bits(u))
    if n <= 1025:  # and the interpreter (slow in Python)
        ienv = dict(n=n, xs=xs, ks=ks, of=np.zeros(n), oi=np.zeros(n, np.int64), scale=7.25, m=11)
        H.run_interp(H.site("map_chain2"), ienv)
        assert np.array_equal(bits(ienv["of"]), bits(u)) and np.array_equal(ienv["oi"], g2)


@pytest.mark.parametrize("n", SIZES)
def test_dot(hyd, orc, prog, arrays, n):
    xs = orc.splitmix_f64(n, 911 + n, symmetric=True)
    ys = orc.splitmix_f64(n, 912 + n, symmetric=True)
    env = dict(acc=0.5, red_dim=n, xs=arrays.host(xs), ys=arrays.host(ys))
    got = call(hyd, prog, H.site("dot"), env)
    exact = orc.dot_dd(xs, ys, init=0.5)
    assert abs(got - exact) <= REL_TOL_REDUCTION * max(abs(exact), 1e-300)
    assert call(hyd, prog, H.site("dot"), env) == got  # deterministic
    if n == 0:
        assert got == 0.5


def test_dot_full_size(hyd, orc, prog, arrays):
    c = workloads.DOTMAP
    n = c["n"]
    xs = orc.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = orc.splitmix_f64(n, c["seed"], n, symmetric=True)
    print("generated dot n=%d" % n, flush=True)
    got = call(hyd, prog, H.site("dot"), dict(acc=0.0, red_dim=n, xs=arrays.host(xs), ys=arrays.host(ys)))
    exact = orc.dot_dd(xs, ys)
    assert abs(got - exact) <= REL_TOL_REDUCTION * abs(exact)
    ones = arrays.host(np.ones(n))
    assert call(hyd, prog, H.site("dot"), dict(acc=0.0, red_dim=n, xs=ones, ys=ones)) == float(n)


@pytest.mark.parametrize("n", [0, 1, 777, 1025, 300_007])
def test_integer_and_product_reductions(hyd, prog, arrays, n):
    ks = np.random.default_rng(n).integers(-(1 << 62), 1 << 62, n, dtype=np.int64)
    dk = arrays.host(ks)
    got = call(hyd, prog, H.site("reduce_i64", 0), dict(acc=5, ks=dk, n=n))
    with np.errstate(over="ignore"):
        sq = (ks.astype(np.uint64) * ks.astype(np.uint64)).sum(dtype=np.uint64) + np.uint64(5)
    assert got == int(np.int64(sq.astype(np.int64)))
    got = call(hyd, prog, H.site("reduce_i64", 1), dict(prod=1, ks=dk, n=n))
    f = (np.fmod(ks, 5) + 3).astype(np.uint64)
    p = np.uint64(1)
    with np.errstate(over="ignore"):
        for x in f:  # exact mod 2^64
            p = p * x
    assert got == int(np.array([p]).astype(np.int64)[0])
    xs = np.random.default_rng(n + 1).uniform(-1, 1, n)
    got = call(hyd, prog, H.site("reduce_i64", 2), dict(n=n, pf=1.0, xs=arrays.host(xs)))
    ref = float(np.prod(np.longdouble(1.0) + np.longdouble(xs * 1.0e-3))) if n else 1.0
    assert abs(got - ref) <= REL_TOL_REDUCTION * abs(ref)


def run_jacobi(hyd, prog, arrays, u_arr, out_arr, shape):
    h, w = shape
    env = dict(u=u_arr, out=out_arr, shp=(h, w), n0=h, n1=w, n0m1=h - 1, n1m1=w - 1, bd_start_r=h - 1,
               bd_start_c=w - 1)
    for s in H.SITES["jacobi_step"]:
        call(hyd, prog, s, env)


@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (64, 130), (257, 1000), (1000, 63)])
def test_jacobi_regions_one_step(hyd, orc, prog, arrays, shape):
    u = orc.splitmix_f64(shape[0] * shape[1], 931 + shape[1]).reshape(shape)
    out = arrays.device(shape)
    run_jacobi(hyd, prog, arrays, arrays.host(u), out, shape)
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(orc.stencil2d(u, 0, 0.1)))


def test_jacobi_regions_iterate(hyd, orc, prog, arrays):
    # the iterate's ping-pong on the host around device-resident buffers
    shape, steps = (700, 513), 7
    u = orc.splitmix_f64(shape[0] * shape[1], 933).reshape(shape)
    a, b = arrays.host(u), arrays.device(shape)
    for _ in range(steps):
        run_jacobi(hyd, prog, arrays, a, b, shape)
        a, b = b, a
    assert np.array_equal(bits(arrays.read(a, np.float64)), bits(orc.iterate_stencil2d(u, steps, 0, 0.1)))


def test_jacobi_regions_config_grid(hyd, orc, prog, arrays):
    c = workloads.JACOBI
    h, w = c["h"], c["w"]
    u = orc.splitmix_f64(h * w, c["seed"]).reshape(h, w)
    out = arrays.device((h, w))
    print("generated jacobi step %dx%d" % (h, w), flush=True)
    run_jacobi(hyd, prog, arrays, arrays.host(u), out, (h, w))
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(orc.stencil2d(u, 0, 0.1)))


def jacobi_env(u_arr, out_arr, shape):
    h, w = shape
    return dict(u=u_arr, out=out_arr, shp=(h, w), n0=h, n1=w, n0m1=h - 1, n1m1=w - 1, bd_start_r=h - 1,
                bd_start_c=w - 1)


@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (64, 130), (257, 1000), (1000, 63), (8192, 8192)])
def test_jacobi_group_one_launch(hyd, orc, prog, arrays, shape):
    # the 9 regions through their group launcher: one launch (their writes
    # are disjoint boxes of out, and none reads out), the same bits
    print("jacobi group %dx%d" % shape, flush=True)
    u = orc.splitmix_f64(shape[0] * shape[1], 977 + shape[1]).reshape(shape)
    out = arrays.device(shape)
    prog.call("hydk_jacobi_step", jacobi_env(arrays.host(u), out, shape))
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(orc.stencil2d(u, 0, 0.1)))


def test_jacobi_group_iterate(hyd, orc, prog, arrays):
    shape, steps = (700, 513), 9
    u = orc.splitmix_f64(shape[0] * shape[1], 979).reshape(shape)
    a, b = arrays.host(u), arrays.device(shape)
    for _ in range(steps):
        prog.call("hydk_jacobi_step", jacobi_env(a, b, shape))
        a, b = b, a
    assert np.array_equal(bits(arrays.read(a, np.float64)), bits(orc.iterate_stencil2d(u, steps, 0, 0.1)))


def run_iterate(prog, arrays, name, u, steps):
    """<name>(steps, u) as the printed C calls it: the result (the caller's to
    free), whether it is the input array itself, and which form ran (0: no
    step, 1: the loops' launchers in order, 2: temporal)."""
    import ctypes

    inp = arrays.host(u)
    res = getattr(prog, name)(steps, inp)
    same = ctypes.addressof(res.contents) == ctypes.addressof(inp.contents)
    if not same:
        arrays.made.append(res)
    form = getattr(prog._cdll, name + "__form")()
    return arrays.read(res, u.dtype), same, form


def run_heat_iterate(prog, arrays, u, steps):
    return run_iterate(prog, arrays, "hydk_heat_iterate", u, steps)


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (7, 1), (2, 2), (3, 5), (33, 65), (37, 91), (130, 70), (257, 1031)])
@pytest.mark.parametrize("steps", [0, 1, 7, 8, 9, 17])
def test_heat_iterate_generated(hyd, orc, prog, arrays, shape, steps):
    # the whole iterate in one launcher: grids of at least 2 x 2 take the
    # temporal kernel (8 steps per launch, tiles with a halo), smaller ones the
    # small-grid loop's launcher step by step; zero steps return the input
    print("heat iterate %dx%d steps=%d" % (shape + (steps,)), flush=True)
    u = orc.splitmix_f64(shape[0] * shape[1], 4001 + 31 * shape[0] + shape[1]).reshape(shape)
    got, same, form = run_heat_iterate(prog, arrays, u, steps)
    assert same == (steps == 0)
    assert form == (0 if steps == 0 else 2 if min(shape) >= 2 else 1)
    assert np.array_equal(bits(got.reshape(shape)), bits(orc.iterate_stencil2d(u, steps, 0, 0.1)))


def test_heat_iterate_generated_config_grid(hyd, orc, prog, arrays):
    c = workloads.JACOBI
    h, w, steps = c["h"], c["w"], 20
    print("generated heat iterate %dx%d steps=%d" % (h, w, steps), flush=True)
    u = orc.splitmix_f64(h * w, c["seed"]).reshape(h, w)
    got, same, form = run_heat_iterate(prog, arrays, u, steps)
    assert not same and form == 2
    assert np.array_equal(bits(got.reshape(h, w)), bits(orc.iterate_stencil2d(u, steps, 0, c["alpha"])))


@pytest.mark.parametrize("shape", [(1, 1), (2, 7), (3, 3), (4, 4), (5, 9), (33, 65), (130, 70), (257, 1031)])
@pytest.mark.parametrize("steps", [0, 1, 3, 4, 5, 9])
def test_wide_iterate_generated(hyd, prog, arrays, shape, steps):
    # radius 2: 4 steps per launch, 25 region loops in the temporal run from
    # 4 x 4 up, the small-grid loop below; against numpy's restatement
    print("wide iterate %dx%d steps=%d" % (shape + (steps,)), flush=True)
    u = np.random.default_rng(7 * shape[0] + shape[1] + steps).uniform(0, 1, shape)
    got, same, form = run_iterate(prog, arrays, "hydk_wide_iterate", u, steps)
    assert same == (steps == 0)
    assert form == (0 if steps == 0 else 2 if min(shape) >= 4 else 1)
    assert np.array_equal(bits(got.reshape(shape)), bits(H.wide_iterate_numpy(u, steps)))


def test_wide_iterate_generated_large(hyd, prog, arrays):
    shape, steps = (2048, 3000), 9
    print("wide iterate %dx%d steps=%d" % (shape + (steps,)), flush=True)
    u = np.random.default_rng(99).uniform(0, 1, shape)
    got, same, form = run_iterate(prog, arrays, "hydk_wide_iterate", u, steps)
    assert form == 2
    assert np.array_equal(bits(got.reshape(shape)), bits(H.wide_iterate_numpy(u, steps)))


@pytest.mark.parametrize("shape,steps", [((1, 1), 2), ((1, 9), 1), ((2, 2), 9), ((37, 91), 9), ((513, 1031), 17),
                                         ((2048, 2000), 10)])
def test_int_iterate_generated(hyd, prog, arrays, shape, steps):
    # int64 state through the temporal form (from 2 x 2 up): wrapping sums and C %
    print("int iterate %dx%d steps=%d" % (shape + (steps,)), flush=True)
    u = np.random.default_rng(shape[1] + steps).integers(0, 1 << 40, shape, dtype=np.int64)
    got, same, form = run_iterate(prog, arrays, "hydk_int_iterate", u, steps)
    assert form == (2 if min(shape) >= 2 else 1)
    assert np.array_equal(got.reshape(shape), H.int_iterate_numpy(u, steps))


@pytest.mark.parametrize("shape,steps", [((1, 1), 2), ((3, 3), 3), ((37, 91), 9), ((513, 1031), 17), ((2048, 2048), 5)])
def test_interior_iterate_falls_back_in_order(hyd, prog, arrays, shape, steps):
    # the loop's box leaves the boundary unwritten: the launcher's tiling check
    # rejects the temporal form, and each step runs into a zeroed array
    print("interior iterate %dx%d steps=%d" % (shape + (steps,)), flush=True)
    u = np.random.default_rng(shape[0] + steps).uniform(0, 1, shape)
    got, same, form = run_iterate(prog, arrays, "hydk_interior_iterate", u, steps)
    assert not same and form == 1
    assert np.array_equal(bits(got.reshape(shape)), bits(H.interior_iterate_numpy(u, steps)))


@pytest.mark.parametrize("n,sh", [(1000, 10), (1000, 0), (1000, 1000), (5, 7), (100_003, 517)])
def test_group_in_order_when_dependent(hyd, prog, arrays, n, sh):
    # out = xs + 1; out[sh:] = 2 xs[sh:] (overlapping writes); ys = out + xs
    # (reads what the others write): the group launcher must run them in order
    xs = np.random.default_rng(n + sh).uniform(-1, 1, n)
    out, ys = arrays.device((n,)), arrays.device((n,))
    prog.call("hydk_overlap", dict(n=n, out=out, xs=arrays.host(xs), sh=sh, ys=ys))
    ref = xs + 1.0
    if sh < n:
        ref[sh:] = xs[sh:] * 2.0
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(ref))
    assert np.array_equal(bits(arrays.read(ys, np.float64)), bits(ref + xs))


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (5, 4), (33, 65), (512, 300)])
def test_wrap_stencil(hyd, prog, arrays, shape):
    u = np.random.default_rng(shape[1]).uniform(0, 1, shape)
    out = arrays.device(shape)
    call(hyd, prog, H.site("wrap_stencil"), dict(n=u.size, out=out, shp=shape, u=arrays.host(u)))
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(H.wrap_numpy(u)))


@pytest.mark.parametrize("n", [3, 4, 1025, 100_001])
def test_mirror_1d(hyd, prog, arrays, n):
    v = np.random.default_rng(n).uniform(-1, 1, n)
    out = arrays.device((n,))
    call(hyd, prog, H.site("mirror_1d"), dict(n=n, out=out, v=arrays.host(v)))
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(H.mirror_numpy(v)))


@pytest.mark.parametrize("n", [0, 1, 5000, 1_000_003])
@pytest.mark.parametrize("nb", [1, 7, 64, 4096, 8192, 65536])
def test_histogram(hyd, orc, prog, arrays, n, nb):
    data = orc.splitmix_i64(n, 941 + n + nb, bits=31)
    if n > 10:
        data[::13] = -data[::13]  # negative remainders: routes outside the bins, dropped
    base = np.random.default_rng(nb).integers(-1000, 1000, nb, dtype=np.int64)
    hist = arrays.host(base)  # scatter accumulates into the defaults
    call(hyd, prog, H.site("histogram"), dict(data=arrays.host(data), hist=hist, n=n, nb=nb))
    # C remainders: a negative x lands in no bin unless x % nb == 0
    with np.errstate(over="ignore"):
        ref = (base.astype(np.uint64) + orc.weighted_histogram(data, nb).astype(np.uint64)).astype(np.int64)
    assert np.array_equal(arrays.read(hist, np.int64), ref)


def test_histogram_full_size(hyd, orc, prog, arrays):
    c = workloads.HIST
    data = orc.splitmix_i64(c["n"], c["seed"], bits=c["bits"])
    hist = arrays.device((c["bins"],))
    call(hyd, prog, H.site("histogram"), dict(data=arrays.host(data), hist=hist, n=c["n"], nb=c["bins"]))
    assert np.array_equal(arrays.read(hist, np.int64), orc.weighted_histogram(data, c["bins"]))


@pytest.mark.parametrize("n,nvox", [(1, 5), (20_000, 997), (1_000_000, 4096), (300_000, 300_000)])
def test_splat_f64_order_exact(hyd, orc, prog, arrays, n, nvox):
    idx, val, keep = H.splat_inputs(n, nvox)
    base = np.random.default_rng(nvox).uniform(0, 1, nvox)
    dst = arrays.host(base)
    call(hyd, prog, H.site("splat_f64"), dict(dst=dst, n=n, nvox=nvox))
    ref = orc.scatter_guarded_add(idx, val, keep, base)
    assert np.array_equal(bits(arrays.read(dst, np.float64)), bits(ref))


@pytest.mark.parametrize("shape", [(1, 1), (37, 53), (4096, 300)])
def test_row_sums_serial_order(hyd, prog, arrays, shape):
    m = np.random.default_rng(shape[0]).uniform(-1, 1, shape)
    out = arrays.device((shape[0],))
    call(hyd, prog, H.site("row_sums"), dict(m=arrays.host(m), out=out, out_n=shape[0], red_dim=shape[1]))
    # each row folded left to right from 0.0 by one thread, as the C loop does
    ref = np.cumsum(np.concatenate([np.zeros((shape[0], 1)), m], axis=1), axis=1)[:, -1]
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(ref))


def test_gather_and_map3d(hyd, prog, arrays):
    n, m = 1_000_003, 70_001
    src = np.random.default_rng(1).integers(-(1 << 50), 1 << 50, m, dtype=np.int64)
    perm = np.random.default_rng(2).integers(0, m, n, dtype=np.int64)
    out = arrays.device((n,))
    call(hyd, prog, H.site("gather"), dict(n=n, out=out, perm=arrays.host(perm), src=arrays.host(src)))
    assert np.array_equal(arrays.read(out, np.int64), src[perm])
    s3 = np.random.default_rng(3).uniform(-1, 1, (17, 33, 65))
    out3 = arrays.device(s3.shape)
    call(hyd, prog, H.site("map3d"), dict(n0=17, n1=33, n2=65, out=out3, shp=s3.shape, src=arrays.host(s3)))
    assert np.array_equal(bits(arrays.read(out3, np.float64)), bits(H.map3d_numpy(s3)))


def test_chained_sites_stay_in_hbm(hyd, orc, prog, arrays):
    # map chain -> dot over its output: the second launcher reads the first's
    # result in place (no host copy in between), the host copy refreshed only
    # when asked
    n = 200_003
    xs = orc.splitmix_f64(n, 951, symmetric=True)
    ys = orc.splitmix_f64(n, 952, symmetric=True)
    out = arrays.device((n,))
    call(hyd, prog, H.site("map_chain"), dict(n=n, out=out, xs=arrays.host(xs), ys=arrays.host(ys)))
    L = hyd.lib()
    assert L.hyd_array_is_device_resident(out) == 1
    got = call(hyd, prog, H.site("dot"), dict(acc=0.0, red_dim=n, xs=out, ys=out))
    m = orc.map_chain(xs, ys)
    exact = orc.dot_dd(m, m)
    assert abs(got - exact) <= REL_TOL_REDUCTION * abs(exact)
    assert L.hyd_array_is_device_resident(out) == 1
    assert np.array_equal(bits(arrays.read(out, np.float64)), bits(m))
