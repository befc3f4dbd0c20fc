"""Benchmark of the fused array kernels on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only NAME]

Headline workload (one "step"): the fp64 dot product and the fused 8-op map
chain over the same two vectors of 2^27 elements, uniform in [-1, 1), inputs
resident in HBM.  value = algorithmic bytes of the step (dot: read 2 x 1 GiB;
map chain: read 2 x 1 GiB + write 1 GiB) / wall time per step, in GB/s.
The roofline object is the map-chain kernel's (the step's dominant kernel),
timed with HIP events on the stream it runs on.  cpu_baseline is the CPU
oracle (a C restatement of the emitted OpenMP C) on the same workload, on
all host cores.  The other fused kernels are measured in the same run and
reported under "kernels"; unless --no-cpu, each one also carries the
oracle's time on the host for the same inputs (timing.h style: warmup, then
timed runs, min and mean; bounded samples where the CPU would take minutes)
and whether this run's GPU result matches the oracle's ("matches_oracle").
"dropin_host_arrays" is the same step through the drop-in exports with the
arrays on the host, PCIe transfers included (a reported figure, not value).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydrangea-compiler_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
# The CPU oracle's OpenMP threads are left unbound: with OMP_PROC_BIND set,
# libgomp pins the process's main thread to one core once the oracle runs,
# and every thread started after that (the serving workers, the parallel file
# reads) inherits the one-core mask.

import hydhip  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel_prefix, elements):
    """HBM bytes per launch of a kernel as recorded by the rocprofv3 counter
    passes in profiles/hbm_traffic.json (FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the copy-kernel calibration), or None when that
    record is absent or was taken on a workload of another size."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    for name, k in rec.get("kernels", {}).items():
        if name.startswith("void " + kernel_prefix) or name.startswith(kernel_prefix):
            return k["traffic_bytes"] if k.get("elements") == elements else None
    return None


def timed_loop(fn, steps, warmup, events=None):
    """Runs fn() warmup times, then `steps` times between device syncs.
    Returns (ms_per_step, list of per-step event times in ms if events)."""
    for _ in range(warmup):
        fn(None)
    hydhip.device_sync()
    pairs = [(hydhip.Event(), hydhip.Event()) for _ in range(steps)] if events else None
    t0 = time.perf_counter()
    for k in range(steps):
        fn(pairs[k] if pairs else None)
    hydhip.device_sync()
    t1 = time.perf_counter()
    ev = [a.elapsed_ms(b) for a, b in pairs] if pairs else []
    return (t1 - t0) * 1e3 / steps, ev


def cpu_time(fn, warmup=3, iters=10):
    """The oracle on the host (all cores), timed like the reference's timing.h:
    `warmup` untimed runs, then `iters` timed ones.  Returns (timing, result)."""
    r = None
    for _ in range(warmup):
        r = fn()
    ts = []
    for _ in range(iters):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return dict(min_ms=min(ts) * 1e3, mean_ms=float(np.mean(ts)) * 1e3, warmup=warmup, iters=iters), r


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float64:
        return bool(a.shape == b.shape and (a.view(np.uint64) == b.view(np.uint64)).all())
    return bool(np.array_equal(a, b))


def bench_dotmap(steps, warmup, cpu=True):
    c = workloads.DOTMAP
    n = c["n"]
    xs = hydhip.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = hydhip.splitmix_f64(n, c["seed"], n, symmetric=True)
    out = hydhip.DeviceArray((n,), np.float64)
    dres = hydhip.DeviceArray((1,), np.float64)
    hydhip.device_sync()

    def step(ev):
        hydhip.dot(xs, ys, out=dres)
        if ev:
            ev[0].record()
        hydhip.map_chain(xs, ys, out=out)
        if ev:
            ev[1].record()

    ms, ev = timed_loop(step, steps, warmup, events=True)
    map_ms = float(np.mean(ev))

    # dot alone, for its own line
    def dot_only(evp):
        if evp:
            evp[0].record()
        hydhip.dot(xs, ys, out=dres)
        if evp:
            evp[1].record()

    _, dev_ev = timed_loop(dot_only, steps, 1, events=True)
    dot_ms = float(np.mean(dev_ev))
    dot_bytes = 2 * 8 * n
    map_bytes = 3 * 8 * n
    res = dict(
        n=n, ms_per_step=ms, step_bytes=dot_bytes + map_bytes,
        map_chain=dict(ms=map_ms, bytes=map_bytes, gbs=map_bytes / map_ms / 1e6),
        dot=dict(ms=dot_ms, bytes=dot_bytes, gbs=dot_bytes / dot_ms / 1e6),
        result=float(dres.numpy()[0]),
    )
    if cpu:
        # the same run's results against the oracle: map chain bit for bit,
        # dot within 1e-12 of a double-double sum
        import pyoracle

        hx = pyoracle.splitmix_f64(n, c["seed"], 0, symmetric=True)
        hy = pyoracle.splitmix_f64(n, c["seed"], n, symmetric=True)
        ref = pyoracle.dot_dd(hx, hy)
        res["dot"]["matches_oracle"] = abs(res["result"] - ref) <= 1e-12 * max(1.0, abs(ref))
        res["map_chain"]["matches_oracle"] = bits_equal(out.numpy(), pyoracle.map_chain(hx, hy))
    for a in (xs, ys, out):
        a.free()
    return res


def bench_copy(steps, warmup):
    """Plain streaming copy of 2^27 fp64 (1 GiB read + 1 GiB write): the
    bandwidth the fused kernels are compared with."""
    n = 1 << 27
    a = hydhip.splitmix_f64(n, 7)
    b = hydhip.DeviceArray((n,), np.float64)
    lib = hydhip.lib()

    def step(ev):
        if ev:
            ev[0].record()
        hydhip._check(lib.hydhip_copy_f64(a.ptr, n, b.ptr, None))
        if ev:
            ev[1].record()

    _, ev = timed_loop(step, steps, warmup, events=True)
    ms = float(np.mean(ev))
    a.free()
    b.free()
    return dict(ms=ms, bytes=16 * n, gbs=16 * n / ms / 1e6)


def bench_shapes(steps, warmup):
    """The streaming grid of the map chain and the copy over 2^27 fp64 in the
    flat form and at 1, 2, 4 and 8 workgroups per CU (set by hand, then
    removed), and what the
    autotuner picks per size class.  Map chain and copy results do not depend on
    the grid; the dot keeps its fixed grid (its partial count fixes the order of
    its sum) and is not swept."""
    n = 1 << 27
    c = workloads.DOTMAP
    L = hydhip.lib()
    xs = hydhip.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = hydhip.splitmix_f64(n, c["seed"], n, symmetric=True)
    out = hydhip.DeviceArray((n,), np.float64)
    res = {}
    for kind, name in ((hydhip.SHAPE_MAP_CHAIN, "map_chain"), (hydhip.SHAPE_COPY, "copy")):
        row = {}
        for pc in (hydhip.SHAPE_FLAT, 1, 2, 4, 8):
            hydhip.tune_shape_set(kind, n, pc)

            def step(ev):
                if ev:
                    ev[0].record()
                if kind == hydhip.SHAPE_MAP_CHAIN:
                    hydhip._check(L.hydhip_map_chain_f64(xs.ptr, ys.ptr, n, out.ptr, None))
                else:
                    hydhip._check(L.hydhip_copy_f64(xs.ptr, n, out.ptr, None))
                if ev:
                    ev[1].record()

            _, ev = timed_loop(step, steps, warmup, events=True)
            row["flat" if pc == hydhip.SHAPE_FLAT else str(pc)] = round(float(np.mean(ev)) * 1e3, 1)
            hydhip.tune_shape_set(kind, n, 0)
        best = min(row.values())
        res[name] = dict(us_by_workgroups_per_cu=row, default="flat", default_vs_best=round(row["flat"] / best, 4))
    t0 = time.perf_counter()
    classes = hydhip.tune_shapes()
    res["tuner"] = dict(classes=classes, seconds=round(time.perf_counter() - t0, 3),
                        table=[dict(kind=["map_chain", "copy"][k], n_log2=nl, per_cu=pc)
                               for k, nl, pc in hydhip.tune_shape_table()])
    for k, nl, _ in hydhip.tune_shape_table():  # leave the defaults for the rest of the run
        hydhip.tune_shape_set(k, 1 << nl, 0)
    xs.free()
    ys.free()
    out.free()
    return res


def bench_dropin_host(reps=4):
    """The dot + map chain step through the drop-in exports with host-resident
    inputs (hyd_export_dot, hyd_export_map_chain; host sync on): each call
    uploads its inputs and the map chain's result comes back, so the rate
    includes 5 GB over PCIe per step.  A reported figure, never `value`."""
    c = workloads.DOTMAP
    n = c["n"]
    lib = hydhip.lib()
    prev = lib.hyd_hip_get_host_sync()
    lib.hyd_hip_set_host_sync(1)
    dx = hydhip.splitmix_f64(n, c["seed"], 0, symmetric=True)
    dy = hydhip.splitmix_f64(n, c["seed"], n, symmetric=True)
    xs, ys = dx.numpy(), dy.numpy()
    dx.free()
    dy.free()
    ax, ay = hydhip.host_array(xs), hydhip.host_array(ys)
    del xs, ys
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        lib.hyd_export_dot(ax, ay)
        m = lib.hyd_export_map_chain(ax, ay)
        dt = time.perf_counter() - t0
        lib.hyd_array_free(m)
        if r:  # the first call sets up the pinned halves
            times.append(dt)
    for a in (ax, ay):
        lib.hyd_array_free(a)
    lib.hyd_hip_set_host_sync(prev)
    ms = 1e3 * min(times)
    step_bytes = 40 * n
    return dict(ms_per_step=ms, algorithmic_gbs=step_bytes / ms / 1e6, pcie_bytes_per_step=40 * n,
                note="host inputs uploaded on every call (2 x 2 GiB), result downloaded (1 GiB)")


def bench_guarded(steps, warmup, cpu=True):
    """Guarded scatters on the reference's voxel benchmark shapes: int64 (+)
    rasterization (5M points) and the order-exact fp64 (+.) trilinear splat
    (1M points = 8M contributions), both into 64^3 voxels.  Algorithmic bytes:
    idx + vals + guard per element read, the voxel grid read and written."""
    res = {}
    for name, gen, cfg, dtype in (("voxel_rasterization_i64", workloads.voxel_rasterization, workloads.VOXEL,
                                   np.int64),
                                  ("voxel_trilinear_splat_f64", workloads.voxel_trilinear_splat,
                                   workloads.SPLAT, np.float64)):
        idx, vals, guard, nvox = gen(**cfg)
        di = hydhip.DeviceArray.from_numpy(idx)
        dv = hydhip.DeviceArray.from_numpy(vals.astype(dtype))
        dg = hydhip.DeviceArray.from_numpy(guard)
        dst = hydhip.DeviceArray((nvox,), dtype)
        zeros = hydhip.DeviceArray.from_numpy(np.zeros(nvox, dtype))

        def step(ev):
            hydhip.copy_device(zeros, dst)  # fill [nvox] 0
            if ev:
                ev[0].record()
            hydhip.scatter_guarded_add(di, dv, dg, dst)
            if ev:
                ev[1].record()

        ms, ev = timed_loop(step, steps, warmup, events=True)
        kms = float(np.mean(ev))
        b = 24 * idx.size + 16 * nvox
        res[name] = dict(elements=int(idx.size), kept=int(guard.sum()), voxels=nvox, ms=kms,
                         bytes=b, gbs=b / kms / 1e6, melem_s=idx.size / kms / 1e3)
        if cpu:
            import pyoracle

            zh = np.zeros(nvox, dtype)
            t, ref = cpu_time(lambda: pyoracle.scatter_guarded_add(idx, vals.astype(dtype), guard, zh),
                              warmup=1, iters=3)
            res[name]["matches_oracle"] = bits_equal(dst.numpy(), ref)
            res[name]["cpu_oracle"] = dict(t, threads=1, note="the serial loop the reference emits")
        for a in (di, dv, dg, dst, zeros):
            a.free()
    return res


def _device_array(shape, fill=None):
    """A hyd_array_t* in HBM (zero, or filled on the device by fill(ptr, n));
    its HBM copy is the current one."""
    L = hydhip.lib()
    arr = L.hyd_array_alloc_device8(hydhip.hyd_tuple(*shape))
    if fill is not None:
        n = int(np.prod(shape))
        fill(L.hyd_array_device_data_keep(arr), n)
        L.hyd_array_mark_device_written(arr)
    return arr


def _host_copy(arr, dtype):
    L = hydhip.lib()
    hydhip._check(L.hyd_array_sync_host(arr))
    return hydhip.read_array(arr, dtype)


def bench_generated(steps, warmup, cpu=True):
    """The kernels hydgen generates from the example programs' offloaded
    loops (lib/libhydgen_programs.so), on the workloads the hand-written
    kernels run: the 8-op map chain and the dot over 2^27 fp64 (the dot's
    launcher returns the value, so its time includes the 8-byte read-back),
    the 64-bin weighted histogram over 2^24 int64, one clamp heat step on the
    8192 x 8192 grid as the reference splits it (9 region loops, through their
    group launcher, and one launcher per region), and the 200-step iterate of
    that step through its one launcher.  Each is checked against the oracle in
    the same run (the iterate at 20 steps)."""
    import hydgen_lib

    prog = hydgen_lib.programs()
    L = hydhip.lib()
    res = {}
    c = workloads.DOTMAP
    n = c["n"]

    def fill_f64(seed, counter0):
        return lambda ptr, m: hydhip._check(L.hydhip_fill_splitmix_f64(ptr, m, seed, counter0, 2.0, -1.0, None))

    xs = _device_array((n,), fill_f64(c["seed"], 0))
    ys = _device_array((n,), fill_f64(c["seed"], n))
    out = _device_array((n,))
    made = [xs, ys, out]
    env = dict(n=n, out=out, xs=xs, ys=ys, acc=0.0, red_dim=n)

    def timed(name, nbytes, launchers):
        def step(ev):
            if ev:
                ev[0].record()
            r = None
            for lname in launchers:
                r = prog.call(lname, env)
            if ev:
                ev[1].record()
            return r

        _, ev = timed_loop(step, steps, warmup, events=True)
        ms = float(np.mean(ev))
        res[name] = dict(ms=ms, bytes=nbytes, gbs=nbytes / ms / 1e6, frac_of_8tbs=nbytes / ms / 1e6 / HBM_PEAK_GBS,
                         launchers=len(launchers))
        return step

    timed("map_chain", 24 * n, ["hydk_map_chain_0"])
    dot_step = timed("dot", 16 * n, ["hydk_dot_0"])
    dot_val = dot_step(None)
    if cpu:
        import pyoracle

        hx = pyoracle.splitmix_f64(n, c["seed"], 0, symmetric=True)
        hy = pyoracle.splitmix_f64(n, c["seed"], n, symmetric=True)
        res["map_chain"]["matches_oracle"] = bits_equal(_host_copy(out, np.float64), pyoracle.map_chain(hx, hy))
        ref = pyoracle.dot_dd(hx, hy)
        res["dot"]["matches_oracle"] = bool(abs(dot_val - ref) <= 1e-12 * abs(ref))
        del hx, hy
    h = workloads.HIST
    data = _device_array((h["n"],), lambda ptr, m: hydhip._check(
        L.hydhip_fill_splitmix_i64(ptr, m, h["seed"], 0, 64 - h["bits"], None)))
    hist = _device_array((h["bins"],))
    made += [data, hist]
    env.update(data=data, hist=hist, nb=h["bins"])
    env["n"] = h["n"]
    timed("weighted_histogram", 8 * h["n"] + 8 * h["bins"], ["hydk_histogram_0"])
    if cpu:
        import pyoracle

        hist1 = _device_array((h["bins"],))
        made.append(hist1)
        prog.call("hydk_histogram_0", dict(env, hist=hist1))
        hd = pyoracle.splitmix_i64(h["n"], h["seed"], bits=h["bits"])
        res["weighted_histogram"]["matches_oracle"] = bits_equal(_host_copy(hist1, np.int64),
                                                                 pyoracle.weighted_histogram(hd, h["bins"]))
    j = workloads.JACOBI
    hh, ww = j["h"], j["w"]
    u = _device_array((hh, ww), lambda ptr, m: hydhip._check(L.hydhip_fill_splitmix_f64(ptr, m, j["seed"], 0, 1.0, 0.0,
                                                                                           None)))
    uo = _device_array((hh, ww))
    made += [u, uo]
    env.update(u=u, out=uo, shp=(hh, ww), n0=hh, n1=ww, n0m1=hh - 1, n1m1=ww - 1, bd_start_r=hh - 1,
               bd_start_c=ww - 1)
    # the 9 region loops through their group launcher (one launch: their
    # writes are disjoint boxes of out), and, for comparison, one by one
    regions = sorted(k for k in prog.launchers() if k.startswith("hydk_jacobi_step_"))
    timed("jacobi_heat_step_regions", 16 * hh * ww, regions)
    timed("jacobi_heat_step", 16 * hh * ww, ["hydk_jacobi_step"])
    if cpu:
        import pyoracle

        hu = pyoracle.splitmix_f64(hh * ww, j["seed"]).reshape(hh, ww)
        res["jacobi_heat_step"]["matches_oracle"] = bits_equal(_host_copy(uo, np.float64).reshape(hh, ww),
                                                               pyoracle.stencil2d(hu, 0, j["alpha"]))
    # the whole iterate (200 steps) through one generated launcher: its
    # temporal form, 8 steps per launch; the call allocates the two state
    # buffers it ping-pongs (their allocation is inside the timed call), and
    # the result is freed after the end event
    iters = j["iters"]
    ms_it = []
    form = -1
    for k in range(max(1, warmup) + steps):
        e0, e1 = hydhip.Event(), hydhip.Event()
        e0.record()
        r = prog.hydk_heat_iterate(iters, u)
        e1.record()
        hydhip.device_sync()
        form = prog._cdll.hydk_heat_iterate__form()
        if k >= max(1, warmup):
            ms_it.append(e0.elapsed_ms(e1))
        if k == 0 and cpu:
            import pyoracle

            ref20 = prog.hydk_heat_iterate(20, u)
            res["heat_iterate_check"] = dict(steps=20, matches_oracle=bits_equal(
                _host_copy(ref20, np.float64).reshape(hh, ww), pyoracle.iterate_stencil2d(hu, 20, 0, j["alpha"])))
            L.hyd_array_free(ref20)
        L.hyd_array_free(r)
    ms = float(np.mean(ms_it))
    res["heat_iterate"] = dict(steps_per_call=iters, ms=ms, us_per_step=ms * 1e3 / iters,
                               form="temporal" if form == 2 else "in order" if form == 1 else "none",
                               gbs_algorithmic=16 * hh * ww * iters / ms / 1e6)
    # the radius-2 iterate (9-point cross, 25 region loops; 4 steps per
    # launch) on the same grid, 40 steps per call; its parity is the tests'
    ms_w = []
    for k in range(1 + steps):
        e0, e1 = hydhip.Event(), hydhip.Event()
        e0.record()
        r = prog.hydk_wide_iterate(40, u)
        e1.record()
        hydhip.device_sync()
        form = prog._cdll.hydk_wide_iterate__form()
        if k >= 1:
            ms_w.append(e0.elapsed_ms(e1))
        L.hyd_array_free(r)
    ms = float(np.mean(ms_w))
    res["wide_iterate"] = dict(steps_per_call=40, ms=ms, us_per_step=ms * 1e3 / 40,
                               form="temporal" if form == 2 else "in order" if form == 1 else "none")
    for a in made:
        L.hyd_array_free(a)
    return res


def bench_coo_sort(steps, warmup, cpu=True):
    """COO -> CSR of 1e7 nonzeros over 2^20 x 2^20 (the served mix's largest
    build: rows, cols and vals uniform over 2^20) and sort_indices of 1e7
    40-bit keys (the width of the build's packed keys), standalone.  Rates
    against the compulsory bytes: COO->CSR reads rows, cols, vals and writes
    col_idx, vals (40 B per nonzero) and row_ptr (8 B per row); the sort reads
    the keys and writes the permutation (16 B per key).  Beside them, the
    oracle's serial loops on the same inputs."""
    nnz, nrows = 10_000_000, 1 << 20
    seed = workloads.SERVE["seed"]
    rows = hydhip.splitmix_i64(nnz, seed + 13, bits=20)
    cols = hydhip.splitmix_i64(nnz, seed + 14, bits=20)
    vals = hydhip.splitmix_i64(nnz, seed + 15, bits=20)
    keys = hydhip.splitmix_i64(nnz, seed + 16, bits=40)
    perm = hydhip.DeviceArray((nnz,), np.int64)
    outs = (hydhip.DeviceArray((nrows + 1,), np.int64), hydhip.DeviceArray((nnz,), np.int64),
            hydhip.DeviceArray((nnz,), np.int64))
    got = {}

    def coo(ev):
        # outputs allocated once, outside the timed calls (as the sort's)
        if ev:
            ev[0].record()
        got["coo"] = hydhip.coo_to_csr(nrows, nrows, rows, cols, vals, out=outs)
        if ev:
            ev[1].record()

    def srt(ev):
        if ev:
            ev[0].record()
        hydhip.sort_indices(keys, out=perm)
        if ev:
            ev[1].record()

    scan_out = hydhip.DeviceArray((nnz,), np.int64)

    def scn(ev):
        if ev:
            ev[0].record()
        hydhip.scan(vals, 0, True, False, out=scan_out)
        if ev:
            ev[1].record()

    cms, cev = timed_loop(coo, max(3, steps // 4), 1, events=True)
    sms, sev = timed_loop(srt, max(3, steps // 4), 1, events=True)
    qms, qev = timed_loop(scn, max(3, steps // 2), 1, events=True)
    cb = 40 * nnz + 8 * (nrows + 1)
    res = dict(
        coo_to_csr=dict(nnz=nnz, nrows=nrows, ms_per_call=cms, event_ms=float(np.mean(cev)), compulsory_bytes=cb,
                        gbs=cb / float(np.mean(cev)) / 1e6, merged_nnz=int(got["coo"][3])),
        sort_indices=dict(keys=nnz, key_bits=40, ms_per_call=sms, event_ms=float(np.mean(sev)),
                          compulsory_bytes=16 * nnz, gbs=16 * nnz / float(np.mean(sev)) / 1e6),
        scan_i64=dict(n=nnz, kind="inclusive (+), reduce-then-scan in two launches", ms_per_call=qms,
                      event_ms=float(np.mean(qev)), bytes=16 * nnz, gbs=16 * nnz / float(np.mean(qev)) / 1e6))
    if cpu:
        import pyoracle

        hr, hc, hv, hk = rows.numpy(), cols.numpy(), vals.numpy(), keys.numpy()
        t, ref = cpu_time(lambda: pyoracle.coo_to_csr(nrows, nrows, hr, hc, hv), warmup=0, iters=1)
        rp, ci, cv, m = got["coo"]
        res["coo_to_csr"]["matches_oracle"] = bool(m == ref[3] and bits_equal(rp.numpy(), ref[0]) and
                                                   bits_equal(ci.numpy(), ref[1]) and bits_equal(cv.numpy(), ref[2]))
        res["coo_to_csr"]["cpu_oracle"] = dict(t, threads=1, note="the serial loops the reference emits")
        t, (sref, _) = cpu_time(lambda: pyoracle.sort_indices(hk), warmup=0, iters=1)
        res["sort_indices"]["matches_oracle"] = bits_equal(perm.numpy(), sref)
        res["sort_indices"]["cpu_oracle"] = dict(t, threads=1, note="hyd_radix_sort_entries, serial")
        t, qref = cpu_time(lambda: pyoracle.scan(hv, 0, True, False), warmup=0, iters=1)
        res["scan_i64"]["matches_oracle"] = bits_equal(scan_out.numpy(), qref)
        res["scan_i64"]["cpu_oracle"] = dict(t, threads=1, note="the reference's serial scan loop")
    for a in (rows, cols, vals, keys, perm, scan_out) + outs:
        a.free()
    return res


def csv_file_round_trip(x, n):
    """The drop-in file functions end to end on a device-resident array of n
    fp64 values: hyd_write_array_csv_float (format on the GPU, text back over
    PCIe through pinned buffers, fwrite) and hyd_read_float_array_csv (fread
    into pinned buffers, copy to HBM, parse, copy the values to arr->data as
    the reference's callers expect).  The file lives in /dev/shm when the host
    has it (page cache only, no disk)."""
    import tempfile

    L = hydhip.lib()
    d = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    path = os.path.join(d, "hyd_bench_%d.csv" % os.getpid()).encode()
    arr = L.hyd_array_alloc_device(hydhip.hyd_tuple(n), 8)
    hydhip._check(L.hydhip_memcpy_d2d(L.hyd_array_device_data(arr), x.ptr, 8 * n, None))
    hydhip.device_sync()
    t0 = time.perf_counter()
    L.hyd_write_array_csv_float(arr, path)
    t1 = time.perf_counter()
    back = L.hyd_read_float_array_csv(path, hydhip.hyd_tuple(n))
    t2 = time.perf_counter()
    size = os.path.getsize(path)
    same = bool((hydhip.read_array(back, np.float64).view(np.uint64) == x.numpy().view(np.uint64)).all())
    os.remove(path)
    L.hyd_array_free(arr)
    L.hyd_array_free(back)
    return dict(where=d, file_bytes=size, write_s=t1 - t0, read_s=t2 - t1, write_gbs=size / (t1 - t0) / 1e9,
                read_gbs=size / (t2 - t1) / 1e9, values_back_identical=same)


def bench_csv(steps, warmup, cpu=True):
    """CSV text <-> arrays in HBM, as read_array_float / write_array_float and
    read_array / write_array read and write files (runtime/hydrangea_runtime.h:
    549-594, runtime/hyd_write_csv.c:5-59): the map chain's 2^27-value fp64
    input written as "%.17g" text and read back, and the histogram's 2^24
    int64 input as "%lld" text.  Each call includes its status read-back.
    Algorithmic HBM bytes: format reads 8 B per value and writes the text;
    parse reads the text and writes 8 B per value.  Beside it, the
    reference's loops (glibc fprintf / fscanf, one thread) on 2^20 values."""
    import ctypes

    import pyoracle

    res = {}
    L = hydhip.lib()
    for name, n, is_float in (("f64_2^27", workloads.DOTMAP["n"], True), ("i64_2^24", workloads.HIST["n"], False)):
        seed = 0x435356
        x = hydhip.splitmix_f64(n, seed, 0, symmetric=True) if is_float else hydhip.splitmix_i64(n, seed)
        cap = L.hydhip_csv_text_bound(n)
        text = hydhip.DeviceArray((cap,), np.uint8)
        y = hydhip.DeviceArray((n,), np.float64 if is_float else np.int64)
        nb, ex = ctypes.c_int64(), ctypes.c_int64()
        st = hydhip.CsvStatus()

        def fmt(ev):
            if ev:
                ev[0].record()
            if is_float:
                hydhip._check(L.hydhip_csv_format_f64(x.ptr, n, n, text.ptr, cap, ctypes.byref(nb),
                                                      ctypes.byref(ex), None))
            else:
                hydhip._check(L.hydhip_csv_format_i64(x.ptr, n, n, text.ptr, cap, ctypes.byref(nb), None))
            if ev:
                ev[1].record()

        def parse(ev):
            if ev:
                ev[0].record()
            fn = L.hydhip_csv_parse_f64 if is_float else L.hydhip_csv_parse_i64
            hydhip._check(fn(text.ptr, nb.value, n, y.ptr, ctypes.byref(st), None))
            if ev:
                ev[1].record()

        fms, _ = timed_loop(fmt, steps, warmup, events=True)
        pms, _ = timed_loop(parse, steps, warmup, events=True)
        ok = st.error_kind == 0 and st.numbers_found == n
        t = nb.value
        r = dict(values=n, text_bytes=t, parsed_all=bool(ok), exact_path=int(st.exact_path),
                 format_ms=fms, parse_ms=pms, format_gvalues_s=n / fms / 1e6, parse_gvalues_s=n / pms / 1e6,
                 format_hbm_gbs=(8 * n + t) / fms / 1e6, parse_hbm_gbs=(t + 8 * n) / pms / 1e6)
        if is_float:
            r["file"] = csv_file_round_trip(x, n)
        if cpu:
            k = 1 << 20
            xs = (pyoracle.splitmix_f64(k, seed, 0, symmetric=True) if is_float
                  else pyoracle.splitmix_i64(k, seed))
            t0 = time.perf_counter()
            txt = pyoracle.csv_write(xs, 1, k, is_float)
            t1 = time.perf_counter()
            pyoracle.csv_read(txt, k, is_float)
            t2 = time.perf_counter()
            r["cpu_reference_loop"] = dict(sample="2^20 values, one thread (glibc)",
                                           format_mvalues_s=k / (t1 - t0) / 1e6,
                                           parse_mvalues_s=k / (t2 - t1) / 1e6)
        res[name] = r
        for a in (x, text, y):
            a.free()
    return res


def bench_hist(steps, warmup, cpu=True):
    c = workloads.HIST
    d = hydhip.splitmix_i64(c["n"], c["seed"], bits=c["bits"])
    out = hydhip.DeviceArray((c["bins"],), np.int64)
    res = {}
    for name, strat in (("auto", 0), ("lds", 1), ("global", 2)):
        def step(ev, strat=strat):
            if ev:
                ev[0].record()
            hydhip.weighted_histogram(d, c["bins"], strat, out=out)
            if ev:
                ev[1].record()

        ms, ev = timed_loop(step, steps, warmup, events=True)
        kms = float(np.mean(ev))
        b = 8 * c["n"] + 8 * c["bins"]
        res[name] = dict(ms_per_call=ms, kernel_ms=kms, bytes=b, gbs=b / kms / 1e6)
        if cpu:
            got = out.numpy()
            res[name]["got"] = got
    if cpu:
        import pyoracle

        host = pyoracle.splitmix_i64(c["n"], c["seed"], bits=c["bits"])
        t, ref = cpu_time(lambda: pyoracle.weighted_histogram(host, c["bins"]))
        for name in ("auto", "lds", "global"):
            res[name]["matches_oracle"] = bits_equal(res[name].pop("got"), ref)
        res["cpu_oracle"] = dict(t, threads=pyoracle.num_threads(), gbs=8 * c["n"] / t["mean_ms"] / 1e6)
    d.free()
    return res


def bench_jacobi(steps, warmup, cpu=True):
    c = workloads.JACOBI
    h, w, iters = c["h"], c["w"], c["iters"]
    u = hydhip.splitmix_f64(h * w, c["seed"])
    u.shape = (h, w)
    b0 = hydhip.DeviceArray((h, w), np.float64)
    b1 = hydhip.DeviceArray((h, w), np.float64)

    def run(ev):
        if ev:
            ev[0].record()
        hydhip.iterate_stencil2d(u, iters, buf0=b0, buf1=b1, use_graph=True)
        if ev:
            ev[1].record()

    ms, ev = timed_loop(run, max(1, steps // 4), 1, events=True)
    gms = float(np.mean(ev))
    step_bytes = 2 * 8 * h * w
    res = dict(iters=iters, graph_ms=gms, ms_per_iterate=ms, per_step_us=gms * 1e3 / iters,
               bytes_per_step=step_bytes, gbs=step_bytes * iters / gms / 1e6)
    if cpu:
        # a bounded sample for the CPU: 20 of the 200 steps, the GPU's 20 steps
        # checked bit for bit against the oracle's
        import pyoracle

        k = 20
        host = pyoracle.splitmix_f64(h * w, c["seed"]).reshape(h, w)
        t, ref = cpu_time(lambda: pyoracle.iterate_stencil2d(host, k, alpha=c["alpha"]), warmup=1, iters=3)
        got, _, _ = hydhip.iterate_stencil2d(u, k, alpha=c["alpha"], buf0=b0, buf1=b1, use_graph=True)
        res["matches_oracle_20_steps"] = bits_equal(got.numpy(), ref)
        res["cpu_oracle"] = dict(t, sample="20 steps", threads=pyoracle.num_threads(),
                                 per_step_ms=t["mean_ms"] / k)
    for a in (u, b0, b1):
        a.free()
    return res


def bench_batched(steps, warmup, cpu=True):
    c = workloads.BATCH
    lens, bins = workloads.batch_instances()
    off, hoff = hydhip.batch_offsets(lens, bins)
    total = int(lens.sum())
    d = hydhip.splitmix_i64(total, c["seed"], bits=c["bits"])
    hist = hydhip.DeviceArray((int(bins.sum()),), np.int64)
    t0 = time.perf_counter()
    classes = hydhip.tune_batched()
    tune_s = time.perf_counter() - t0
    res = dict(instances=int(lens.size), elements=total, bytes=workloads.batch_algorithmic_bytes(lens, bins),
               tuned_classes=classes, tune_s=round(tune_s, 3))
    for name, pol in (("tuned", hydhip.BATCH_TUNED), ("lds", hydhip.BATCH_LDS), ("global", hydhip.BATCH_GLOBAL)):
        def step(ev, pol=pol):
            if ev:
                ev[0].record()
            hydhip.batched_histogram(d, off, lens, bins, hoff, hist=hist, policy=pol)
            if ev:
                ev[1].record()

        ms, ev = timed_loop(step, max(3, steps // 4), 1, events=True)
        res[name] = dict(ms_per_batch=ms, event_ms=float(np.mean(ev)),
                         gbs=res["bytes"] / ms / 1e6, req_s=lens.size / ms * 1e3)
    # the instances of one bin class alone (one batched launch per class, in
    # the order of BATCH["bins"]): where the batch's time goes
    by_class = {}
    for nb in c["bins"]:
        m = bins == nb
        if not m.any():
            continue
        lc, bc, oc, hc = lens[m], bins[m], off[m], hoff[m]
        cb = workloads.batch_algorithmic_bytes(lc, bc)

        def step(ev, lc=lc, bc=bc, oc=oc, hc=hc):
            if ev:
                ev[0].record()
            hydhip.batched_histogram(d, oc, lc, bc, hc, hist=hist, policy=hydhip.BATCH_TUNED)
            if ev:
                ev[1].record()

        _, ev = timed_loop(step, max(3, steps // 4), 1, events=True)
        ms = float(np.mean(ev))
        by_class[str(nb)] = dict(instances=int(m.sum()), elements=int(lc.sum()), bytes=cb, event_ms=ms,
                                 gbs=cb / ms / 1e6)
    res["by_class"] = by_class
    # uniform batches over the first 2^27 elements (equal-length instances of
    # one bin count, LDS policy): the streaming rate of single-range tiles
    # against that of range tiles (bins beyond one 16384-word LDS range, each
    # range a workgroup streaming the whole instance through L2)
    uniform = {}
    nu = 1 << 27
    for nb, count in ((4096, 64), (16384, 64), (65536, 64), (65536, 1024)):
        lu = np.full(count, nu // count, np.int64)
        bu = np.full(count, nb, np.int64)
        ou, hu = hydhip.batch_offsets(lu, bu)
        hu_dev = hydhip.DeviceArray((int(bu.sum()),), np.int64)

        def step(ev, lu=lu, bu=bu, ou=ou, hu=hu, hu_dev=hu_dev):
            if ev:
                ev[0].record()
            hydhip.batched_histogram(d, ou, lu, bu, hu, hist=hu_dev, policy=hydhip.BATCH_LDS)
            if ev:
                ev[1].record()

        _, ev = timed_loop(step, max(3, steps // 4), 1, events=True)
        ms = float(np.mean(ev))
        ub = workloads.batch_algorithmic_bytes(lu, bu)
        uniform["%d_x_%d_bins_%d" % (count, nu // count, nb)] = dict(
            event_ms=ms, bytes=ub, gbs=ub / ms / 1e6, element_gbs=8 * nu / ms / 1e6,
            ranges=-(-nb // 16384))
        hu_dev.free()
    res["uniform"] = uniform
    # per-call plain launches: one synchronous hydhip_weighted_histogram_i64 per instance
    lib = hydhip.lib()
    s = hydhip.default_stream()

    def percall(ev):
        for k in range(lens.size):
            rc = lib.hydhip_weighted_histogram_i64(d.ptr + 8 * int(off[k]), int(lens[k]), int(bins[k]),
                                                   hist.ptr + 8 * int(hoff[k]), 0, s)
            if rc:
                raise hydhip.HydError(rc, lib.hydhip_last_error().decode())
            lib.hydhip_stream_sync(s)

    ms, _ = timed_loop(percall, 1, 1)
    res["per_call"] = dict(ms_per_batch=ms, req_s=lens.size / ms * 1e3)
    res["speedup_vs_per_call"] = res["per_call"]["ms_per_batch"] / res["tuned"]["ms_per_batch"]
    if cpu:
        import pyoracle

        host = pyoracle.splitmix_i64(total, c["seed"], bits=c["bits"])
        hist_dev = hist.numpy()
        hydhip.batched_histogram(d, off, lens, bins, hoff, hist=hist, policy=hydhip.BATCH_TUNED)
        t, (ref, _) = cpu_time(lambda: pyoracle.batched_histograms(host, off, lens, bins), warmup=1, iters=3)
        res["matches_oracle"] = bits_equal(hist.numpy(), ref) and bits_equal(hist_dev, ref)
        res["cpu_oracle"] = dict(t, threads=pyoracle.num_threads(), gbs=res["bytes"] / t["mean_ms"] / 1e6)
    d.free()
    hist.free()
    return res


def serve_sample(reqs, per_kind=3):
    """Indices of the served requests checked against the oracle: the largest
    request of each class and the first `per_kind` of each class."""
    size = dict(stencil=lambda r: r["side"] ** 2 * r["iters"], hist=lambda r: r["n"], dot=lambda r: r["n"],
                coo=lambda r: r["nnz"])
    picked = set()
    for kind, key in size.items():
        idx = [k for k, r in enumerate(reqs) if r["kind"] == kind]
        if not idx:
            continue
        picked.add(max(idx, key=lambda k: key(reqs[k])))
        picked.update(idx[:per_kind])
    return picked


def serve_check(reqs, kept, pools):
    """The kept served results against the oracle (integers and stencils bit
    for bit, dots within 1e-12 relative)."""
    import pyoracle

    host = {}

    def pool(name):
        if name not in host:
            host[name] = getattr(pools, name).numpy()
        return host[name]

    ok = {}
    for k, got in sorted(kept.items()):
        r = reqs[k]
        if r["kind"] == "stencil":
            u = pool("grid")[: r["side"] ** 2].reshape(r["side"], r["side"])
            good = bits_equal(got, pyoracle.iterate_stencil2d(u, r["iters"], 0, 0.1).ravel())
        elif r["kind"] == "hist":
            good = bits_equal(got, pyoracle.weighted_histogram(pool("data")[: r["n"]], r["bins"]))
        elif r["kind"] == "dot":
            ref = pyoracle.dot(pool("xs")[: r["n"]], pool("ys")[: r["n"]])
            good = bool(abs(got[0] - ref) <= 1e-12 * abs(ref))
        else:
            n = r["nnz"]
            rp, ci, cv, m = pyoracle.coo_to_csr(r["nrows"], r["ncols"], pool("rows")[:n], pool("cols")[:n],
                                                pool("vals")[:n])
            good = bool(got[3] == m and bits_equal(got[0], rp) and bits_equal(got[1], ci) and bits_equal(got[2], cv))
        ok[k] = good
    return dict(all=all(ok.values()), checked=len(ok), failed=[k for k, v in ok.items() if not v],
                sample="largest request of each class and the first 3 of each class, from an untimed "
                       "all-at-once pass of the same mix",
                requests={str(k): reqs[k] for k in sorted(ok)})


def bench_serve(count=1500, budget=8 << 30, workers=8, load=0.7, cpu=True):
    """Served mixed stream vs plain per-call launches on the same requests.

    plain: one synchronous hydhip_run_request_plain per request (hipMalloc'd
    outputs, plain launches, synchronise, hipFree), back to back.
    served (saturation): every request submitted at t=0; req/s = count / span.
    served (Poisson): arrivals at `load` x the saturation rate; p50/p99 of
    done - submit (server clock)."""
    import servemix

    reqs = workloads.serve_mix(count)
    pools = servemix.Pools(reqs)
    hreqs = [pools.request(r) for r in reqs]
    alg_bytes = sum(workloads.request_bytes_algorithmic(r) for r in reqs)

    # warm both paths (code objects, workspaces)
    for r in hreqs[:8]:
        hydhip.free_plain(hydhip.run_plain(r))

    t0 = time.perf_counter()
    for r in hreqs:
        hydhip.free_plain(hydhip.run_plain(r))
    plain_s = time.perf_counter() - t0

    def drain(srv, outstanding, lat, deadline_s, block_ms):
        got = srv.next_done(block_ms)
        while got is not None:
            rid, res = got
            if res.status != hydhip.HYD_OK:
                raise hydhip.HydError(res.status, "served request failed")
            lat.append(res.done_ms - res.submit_ms)
            srv.release(rid)
            outstanding.discard(rid)
            got = srv.next_done(0.0)

    srv = hydhip.Server(budget, workers)
    try:
        warm = [srv.submit(r) for r in hreqs[:8]]
        for rid in warm:
            srv.wait(rid)
            srv.release(rid)
        while srv.next_done(0.0) is not None:
            pass
        # saturation: everything at once
        lat_sat = []
        outstanding = set()
        t0 = time.perf_counter()
        for r in hreqs:
            outstanding.add(srv.submit(r))
            drain(srv, outstanding, lat_sat, None, 0.0)
        while outstanding:
            drain(srv, outstanding, lat_sat, None, 1000.0)
        sat_s = time.perf_counter() - t0
        st_sat = srv.stats()
        # Poisson arrivals at a fraction of the saturation rate
        rate = load * count / sat_s
        rng = np.random.default_rng(workloads.SERVE["seed"] & 0xFFFFFFFF)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
        lat = []
        outstanding = set()
        t0 = time.perf_counter()
        for k, r in enumerate(hreqs):
            while True:
                wait_s = arrivals[k] - (time.perf_counter() - t0)
                if wait_s <= 0:
                    break
                drain(srv, outstanding, lat, None, min(wait_s * 1e3, 1.0))
            outstanding.add(srv.submit(r))
        while outstanding:
            drain(srv, outstanding, lat, None, 1000.0)
        poisson_s = time.perf_counter() - t0
        st = srv.stats()
        check = None
        if cpu:
            # untimed: the whole mix once more, all at once, with a sample's
            # results kept and compared with the oracle afterwards
            # (results are released as they arrive: outputs held until
            # release count against the HBM budget)
            sample = serve_sample(reqs)
            index = {}
            kept = {}

            def keep(block_ms):
                got = srv.next_done(block_ms)
                while got is not None:
                    rid, res = got
                    if res.status != hydhip.HYD_OK:
                        raise hydhip.HydError(res.status, "served request failed")
                    k = index.pop(rid)
                    if k in sample:
                        kept[k] = servemix.fetch(reqs[k], res)
                    srv.release(rid)
                    got = srv.next_done(0.0)

            for k, r in enumerate(hreqs):
                index[srv.submit(r)] = k
                keep(0.0)
            while index:
                keep(1000.0)
            log("served mix: checking %d sampled results against the oracle" % len(kept))
            check = serve_check(reqs, kept, pools)
    finally:
        srv.close()
        pools.free()
    kinds = {k: sum(1 for r in reqs if r["kind"] == k) for k in ("stencil", "hist", "dot", "coo")}
    return dict(
        requests=count, mix=kinds, hbm_budget=budget, workers=workers,
        algorithmic_bytes=alg_bytes,
        plain=dict(seconds=plain_s, req_s=count / plain_s),
        served_saturation=dict(seconds=sat_s, req_s=count / sat_s, gbs=alg_bytes / sat_s / 1e9,
                               p50_ms=float(np.percentile(lat_sat, 50)),
                               p99_ms=float(np.percentile(lat_sat, 99)),
                               peak_hbm_bytes=st_sat["peak_in_use"], jobs=st_sat["jobs"],
                               batched_jobs=st_sat["batched_jobs"]),
        served_poisson=dict(rate_req_s=rate, seconds=poisson_s, req_s=count / poisson_s,
                            p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99))),
        speedup_vs_plain=plain_s / sat_s,
        totals=dict(completed=st["completed"], failed=st["failed"], rejected=st["rejected"]),
        matches_oracle=None if check is None else check["all"],
        oracle_sample=check,
    )


def cpu_baseline_dotmap():
    import pyoracle

    c = workloads.DOTMAP
    n = c["n"]
    xs = pyoracle.splitmix_f64(n, c["seed"], 0, symmetric=True)
    ys = pyoracle.splitmix_f64(n, c["seed"], n, symmetric=True)
    pyoracle.map_chain(xs[:1 << 20], ys[:1 << 20])
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        pyoracle.dot(xs, ys)
        pyoracle.map_chain(xs, ys)
    dt = (time.perf_counter() - t0) / reps
    return dict(value=5 * 8 * n / dt / 1e9, unit="GB/s", cores=pyoracle.num_threads(), kind="port",
                sample="full workload (2^27 elements), dot + map chain, %d repetitions, %.3f s each" % (reps, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the fused kernels run on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    hydhip.init()
    info = hydhip.device_info()
    log("device", info)
    kernels = {}
    only = set(filter(None, args.only.split(",")))

    log("dot + map chain ...")
    dm = bench_dotmap(args.steps, args.warmup, cpu=not args.no_cpu)
    kernels["dot"] = dm["dot"]
    kernels["map_chain"] = dm["map_chain"]
    if not only or "copy" in only:
        log("copy reference ...")
        kernels["copy_reference"] = bench_copy(args.steps, args.warmup)
    if not only or "dropin" in only:
        log("drop-in with host arrays ...")
        kernels["dropin_host_arrays"] = bench_dropin_host()
    if "shapes" in only:
        log("launch shapes ...")
        kernels["launch_shapes"] = bench_shapes(args.steps, args.warmup)
    if not only or "hist" in only:
        log("histogram ...")
        kernels["weighted_histogram"] = bench_hist(args.steps, args.warmup, cpu=not args.no_cpu)
    if not only or "guarded" in only:
        log("guarded scatters ...")
        kernels["scatter_guarded"] = bench_guarded(args.steps, args.warmup, cpu=not args.no_cpu)
    if not only or "jacobi" in only:
        log("jacobi ...")
        kernels["jacobi_heat"] = bench_jacobi(args.steps, args.warmup, cpu=not args.no_cpu)

    if not only or "generated" in only:
        log("generated kernels ...")
        kernels["generated"] = bench_generated(args.steps, args.warmup, cpu=not args.no_cpu)
    if not only or "coo" in only:
        log("coo -> csr and sort ...")
        kernels["coo_sort"] = bench_coo_sort(args.steps, args.warmup, cpu=not args.no_cpu)

    if not only or "batch" in only:
        log("batched histograms ...")
        kernels["batched_histogram"] = bench_batched(args.steps, args.warmup, cpu=not args.no_cpu)

    if not only or "csv" in only:
        log("csv text ...")
        kernels["csv"] = bench_csv(args.steps, args.warmup, cpu=not args.no_cpu)

    if not only or "serve" in only:
        log("served mix ...")
        kernels["served_mix"] = bench_serve(cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline_dotmap()

    value = dm["step_bytes"] / dm["ms_per_step"] / 1e6
    mc = dm["map_chain"]
    line = {
        "metric": "HBM GB/s & % roofline per fused kernel; batched req/s, p50/p99 latency",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dm["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fp64 dot + fused 8-op map chain over xs, ys [2^27] uniform [-1,1)",
                   "n": dm["n"], "step_bytes": dm["step_bytes"], "device": info["arch"],
                   "cus": info["num_cus"]},
        "roofline": {"bound": "hbm", "kernel": "map_chain_flat_kernel", "achieved": round(mc["gbs"], 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(mc["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("map_chain_flat_kernel<1>", workloads.DOTMAP["n"])},
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
