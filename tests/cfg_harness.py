"""Inputs and the oracle runs of the example programs' offloaded loops (test
infrastructure shared by the CPU and GPU tests of generated kernels)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import cfg_interp  # noqa: E402  (oracle/)
import cfg_programs  # noqa: E402

SITES = {name: cfg_programs.sites(name) for name in cfg_programs.PROGRAMS}


def site(prog, k=0):
    return SITES[prog][k]


def arg_names(s):
    return [v for v, _ in s.get("hsArgs")]


def run_interp(s, env):
    """The oracle: the site's loop run serially over env (arrays in place)."""
    return cfg_interp.run_site(s, env)


def jacobi_env(u):
    h, w = u.shape
    return dict(u=u, out=np.zeros_like(u), shp=(h, w), n0=h, n1=w, n0m1=h - 1, n1m1=w - 1, bd_start_r=h - 1,
                bd_start_c=w - 1)


def jacobi_interp(u):
    """One clamp heat step through the 9 region loops (grids of at least 2 x 2)."""
    env = jacobi_env(np.array(u))
    for s in SITES["jacobi_step"]:
        run_interp(s, env)
    return env["out"]


def splitmix_ints(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, n, dtype=np.int64)


def map_chain2_inputs(n, seed):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(-50.0, 50.0, n)
    xs[::17] = rng.uniform(-1e-3, 1e-3, xs[::17].size)  # tiny divisors
    ks = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    ks[::5] = 0
    ks[1::11] = rng.integers(-(1 << 62), 1 << 62, ks[1::11].size, dtype=np.int64)  # wrapping products
    return xs, ks


def map_chain2_numpy(xs, ks, scale, m):
    """The same chain restated with numpy (C semantics spelled out)."""
    a = xs * scale
    b = np.floor(a)
    c = b.astype(np.int64)
    d = np.fmod(c, m)                                   # C %: sign of the dividend
    e = np.sign(ks) * (np.abs(ks) // 7)                 # C /: toward zero
    with np.errstate(over="ignore"):
        f = d - e
        g = (f.astype(np.uint64) * ks.astype(np.uint64)).astype(np.int64)
    h = d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = h / xs
    q = np.ceil(p)
    r = -q
    lt = xs < 0.0
    s = np.where(lt, r, q)
    both = lt & (ks != 0)
    with np.errstate(over="ignore"):
        g2 = np.where(both, (np.uint64(0) - g.astype(np.uint64)).astype(np.int64), g)
    t = s - 1.5
    u = np.sqrt(np.abs(t))
    return u, g2


WRAP_W = 0.2


def wrap_numpy(u):
    n = np.roll(u, 1, axis=0)    # a (-1) 0 with wrap
    s = np.roll(u, -1, axis=0)   # a 1 0
    w = np.roll(u, 1, axis=1)    # a 0 (-1)
    e = np.roll(u, -1, axis=1)   # a 0 1
    return (0.2 * (n + s) + w * e) - u / 3.0


def wide_step_numpy(u):
    """One step of the radius-2 clamp heat (cfg_programs.wide_iterate) in
    numpy: neighbours by edge padding, the body's IEEE operations in its
    order."""
    h, w = u.shape
    p = np.pad(u, 2, mode="edge")
    a = lambda dr, dc: p[2 + dr:2 + dr + h, 2 + dc:2 + dc + w]  # noqa: E731
    q3 = (a(-1, 0) + a(1, 0)) + (a(0, -1) + a(0, 1))
    q6 = (a(-2, 0) + a(2, 0)) + (a(0, -2) + a(0, 2))
    q10 = (16.0 * q3 - q6) - 60.0 * u
    return u + cfg_programs.WIDE_COEF * q10


def wide_iterate_numpy(u, steps):
    for _ in range(steps):
        u = wide_step_numpy(u)
    return u


def interior_iterate_numpy(u, steps):
    """cfg_programs.interior_iterate: each step a zero array whose interior
    is the heat step of the previous state (the body's operations in order)."""
    for _ in range(steps):
        out = np.zeros_like(u)
        if u.shape[0] > 2 and u.shape[1] > 2:
            c = u[1:-1, 1:-1]
            t3 = ((u[:-2, 1:-1] + u[2:, 1:-1]) + u[1:-1, :-2]) + u[1:-1, 2:]
            out[1:-1, 1:-1] = c + 0.1 * (t3 - 4.0 * c)
        u = out
    return u


def int_iterate_numpy(u, steps):
    """cfg_programs.int_iterate for non-negative int64 grids (numpy's % then
    equals C's): clamp neighbours by edge padding; grids below 2 x 2 become
    zero after a step."""
    for _ in range(steps):
        if u.shape[0] < 2 or u.shape[1] < 2:
            u = np.zeros_like(u)
            continue
        p = np.pad(u, 1, mode="edge")
        h, w = u.shape
        k3 = ((p[0:h, 1:w + 1] + p[2:h + 2, 1:w + 1]) + p[1:h + 1, 0:w]) + p[1:h + 1, 2:w + 2]
        u = (k3 + 3 * u) % 1000003
    return u


def mirror_numpy(v):
    p = np.pad(v, (1, 2), mode="reflect")  # mirror without repeating the edge
    left, centre, right = p[0:-3], p[1:-2], p[3:]
    return (0.5 * left + centre) - 0.25 * right


def splat_inputs(n, nvox):
    i = np.arange(n, dtype=np.int64)
    idx = (i * 17 + 3) % nvox
    val = (i % 13).astype(np.float64) * 0.125 + 0.1
    keep = (i % 3 == 0).astype(np.int64)
    return idx, val, keep


def map3d_numpy(src):
    a, b, c = np.meshgrid(*(np.arange(k, dtype=np.float64) for k in src.shape), indexing="ij")
    return ((a - 2.0 * b) + 0.5 * c) + src
