"""ctypes bindings of oracle/liboracle.so - TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_INT = ctypes.c_int
_D = ctypes.c_double

_SIGS = {
    "oracle_splitmix_i64": (None, [_U64, _U64, _I64, _INT, _P]),
    "oracle_splitmix_f64": (None, [_U64, _U64, _I64, _D, _D, _P]),
    "oracle_num_threads": (_INT, []),
    "oracle_scatter_add_serial": (None, [_P, _P, _I64, _P, _I64]),
    "oracle_scatter_add_privatized": (None, [_P, _P, _I64, _P, _I64]),
    "oracle_scatter_guarded_add_i64": (None, [_P, _P, _P, _I64, _P, _I64]),
    "oracle_scatter_guarded_add_f64": (None, [_P, _P, _P, _I64, _P, _I64]),
    "oracle_scatter_chain_add_f64": (None, [_P, _P, _P, _P, _P, _P, _I64, _P, _I64]),
    "oracle_weighted_histogram": (None, [_P, _I64, _I64, _P]),
    "oracle_weighted_histogram_serial": (None, [_P, _I64, _I64, _P]),
    "oracle_batched_histograms": (None, [_P, _P, _P, _P, _P, _I64, _P]),
    "oracle_dot": (_D, [_P, _P, _I64, _D, _INT]),
    "oracle_dot_serial": (_D, [_P, _P, _I64, _D]),
    "oracle_dot_dd": (_D, [_P, _P, _I64, _D]),
    "oracle_map_chain": (None, [_P, _P, _I64, _P]),
    "oracle_exp": (None, [_P, _I64, _P]),
    "oracle_sqrt": (None, [_P, _I64, _P]),
    "oracle_stencil2d_clamp": (None, [_P, _I64, _I64, _INT, _D, _P]),
    "oracle_iterate_stencil2d": (None, [_P, _I64, _I64, _INT, _D, _I64, _P]),
    "oracle_stencil1d_clamp_sum": (None, [_P, _I64, _P, _INT, _P]),
    "oracle_scan": (None, [_P, _I64, _I64, _INT, _INT, _P]),
    "oracle_segmented_reduce": (None, [_P, _I64, _P, _I64, _P]),
    "oracle_sort_indices": (_INT, [_P, _I64, _P]),
    "oracle_gather": (None, [_P, _I64, _P, _P]),
    "oracle_coo_sum_duplicates": (_I64, [_P, _P, _P, _I64, _I64, _P, _P, _P]),
    "oracle_csr_from_sorted_coo": (None, [_I64, _I64, _P, _P, _P, _I64, _P, _P, _P]),
    "oracle_coo_to_csr": (_I64, [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P]),
    "oracle_csv_read": (_INT, [ctypes.c_char_p, _I64, _I64, _INT, _P, _P]),
    "oracle_csv_write": (_P, [_P, _I64, _I64, _INT, _P]),
    "oracle_free": (None, [_P]),
    "oracle_scan_tokens": (None, [ctypes.c_char_p, _P, _I64, _INT, _P, _P, _P]),
    "oracle_format_values": (None, [_P, _I64, _INT, _P, _I64, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so not built: run oracle/build.py")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data if a.size else None


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads():
    return lib().oracle_num_threads()


def splitmix_i64(n, seed, counter0=0, bits=64):
    out = np.empty(n, np.int64)
    lib().oracle_splitmix_i64(seed & (2**64 - 1), counter0, n, 64 - bits, _p(out))
    return out


def splitmix_f64(n, seed, counter0=0, symmetric=False):
    out = np.empty(n, np.float64)
    scale, offset = (2.0, -1.0) if symmetric else (1.0, 0.0)
    lib().oracle_splitmix_f64(seed & (2**64 - 1), counter0, n, scale, offset, _p(out))
    return out


def scatter_add(idx, vals, dst, privatized=True):
    idx, vals, dst = _i64(idx), _i64(vals), _i64(dst).copy()
    f = lib().oracle_scatter_add_privatized if privatized else lib().oracle_scatter_add_serial
    f(_p(idx), _p(vals), idx.size, _p(dst), dst.size)
    return dst


def scatter_guarded_add(idx, vals, guard, dst):
    """Serial guarded scatter (+) / (+.) into a copy of dst (the defaults);
    guard None: unguarded.  fp64 when vals/dst are float64."""
    idx = _i64(idx)
    g = None if guard is None else _i64(guard)
    if np.asarray(dst).dtype == np.float64:
        vals = np.ascontiguousarray(vals, np.float64)
        out = np.ascontiguousarray(dst, np.float64).copy()
        lib().oracle_scatter_guarded_add_f64(_p(idx), _p(vals), None if g is None else _p(g), idx.size,
                                             _p(out), out.size)
    else:
        vals = _i64(vals)
        out = _i64(dst).copy()
        lib().oracle_scatter_guarded_add_i64(_p(idx), _p(vals), None if g is None else _p(g), idx.size,
                                             _p(out), out.size)
    return out


def scatter_chain_add_f64(phases, dst):
    """phases: list of (idx, vals, guard or None); serial, phase by phase."""
    idx = _i64(np.concatenate([np.asarray(p[0], np.int64) for p in phases]) if phases else np.zeros(0))
    vals = np.ascontiguousarray(np.concatenate([np.asarray(p[1], np.float64) for p in phases])
                                if phases else np.zeros(0), np.float64)
    guard = _i64(np.concatenate([np.asarray(p[2] if p[2] is not None else np.ones(len(p[0])), np.int64)
                                 for p in phases]) if phases else np.zeros(0))
    lens = _i64([len(p[0]) for p in phases])
    off = _i64(np.concatenate([[0], np.cumsum(lens)[:-1]]) if phases else np.zeros(0))
    has = _i64([0 if p[2] is None else 1 for p in phases])
    out = np.ascontiguousarray(dst, np.float64).copy()
    lib().oracle_scatter_chain_add_f64(_p(idx), _p(vals), _p(guard), _p(off), _p(lens), _p(has),
                                       len(phases), _p(out), out.size)
    return out


def weighted_histogram(data, nbins, serial=False):
    data = _i64(data)
    out = np.empty(nbins, np.int64)
    f = lib().oracle_weighted_histogram_serial if serial else lib().oracle_weighted_histogram
    f(_p(data), data.size, nbins, _p(out))
    return out


def batched_histograms(data, off, lens, bins):
    data, off, lens, bins = _i64(data), _i64(off), _i64(lens), _i64(bins)
    hoff = np.zeros(bins.size, np.int64)
    if bins.size > 1:
        hoff[1:] = np.cumsum(bins)[:-1]
    hist = np.empty(int(bins.sum()), np.int64)
    lib().oracle_batched_histograms(_p(data), _p(off), _p(lens), _p(bins), _p(hoff), bins.size, _p(hist))
    return hist, hoff


def dot(xs, ys, init=0.0, nthreads=0):
    xs, ys = _f64(xs), _f64(ys)
    return lib().oracle_dot(_p(xs), _p(ys), min(xs.size, ys.size), init, nthreads)


def dot_serial(xs, ys, init=0.0):
    xs, ys = _f64(xs), _f64(ys)
    return lib().oracle_dot_serial(_p(xs), _p(ys), min(xs.size, ys.size), init)


def dot_dd(xs, ys, init=0.0):
    xs, ys = _f64(xs), _f64(ys)
    return lib().oracle_dot_dd(_p(xs), _p(ys), min(xs.size, ys.size), init)


def map_chain(xs, ys):
    xs, ys = _f64(xs), _f64(ys)
    n = min(xs.size, ys.size)
    out = np.empty(n, np.float64)
    lib().oracle_map_chain(_p(xs), _p(ys), n, _p(out))
    return out


def libm_exp(x):
    x = _f64(x)
    out = np.empty_like(x)
    lib().oracle_exp(_p(x), x.size, _p(out))
    return out


def libm_sqrt(x):
    x = _f64(x)
    out = np.empty_like(x)
    lib().oracle_sqrt(_p(x), x.size, _p(out))
    return out


def stencil2d(u, body=0, alpha=0.1):
    u = _f64(u)
    h, w = u.shape
    out = np.empty_like(u)
    lib().oracle_stencil2d_clamp(_p(u), h, w, body, alpha, _p(out))
    return out


def iterate_stencil2d(u, iters, body=0, alpha=0.1):
    u = _f64(u)
    h, w = u.shape
    out = np.empty_like(u)
    lib().oracle_iterate_stencil2d(_p(u), h, w, body, alpha, iters, _p(out))
    return out


def stencil1d_clamp_sum(a, offsets):
    a = _i64(a)
    offs = _i64(offsets)
    out = np.empty_like(a)
    lib().oracle_stencil1d_clamp_sum(_p(a), a.size, _p(offs), offs.size, _p(out))
    return out


def scan(a, init=0, inclusive=False, reverse=False):
    a = _i64(a)
    out = np.empty_like(a)
    lib().oracle_scan(_p(a), a.size, init, int(inclusive), int(reverse), _p(out))
    return out


def segmented_reduce(offsets, vals, init=0):
    offsets, vals = _i64(offsets), _i64(vals)
    out = np.empty(max(offsets.size - 1, 0), np.int64)
    lib().oracle_segmented_reduce(_p(offsets), offsets.size, _p(vals), init, _p(out))
    return out


def sort_indices(keys):
    keys = _i64(keys)
    out = np.empty_like(keys)
    passes = lib().oracle_sort_indices(_p(keys), keys.size, _p(out))
    return out, passes


def gather(perm, src):
    perm, src = _i64(perm), _i64(src)
    out = np.empty_like(perm)
    lib().oracle_gather(_p(perm), perm.size, _p(src), _p(out))
    return out


def coo_sum_duplicates(rows, cols, vals, nnz=None):
    rows, cols, vals = _i64(rows), _i64(cols), _i64(vals)
    n = rows.size
    nnz = n if nnz is None else nnz
    o = [np.empty(n, np.int64) for _ in range(3)]
    m = lib().oracle_coo_sum_duplicates(_p(rows), _p(cols), _p(vals), nnz, n, _p(o[0]), _p(o[1]), _p(o[2]))
    return o[0], o[1], o[2], m


def csr_from_sorted_coo(nrows, nnz, rows, cols, vals):
    rows, cols, vals = _i64(rows), _i64(cols), _i64(vals)
    n = cols.size
    rp = np.empty(nrows + 1, np.int64)
    ci = np.empty(n, np.int64)
    cv = np.empty(n, np.int64)
    lib().oracle_csr_from_sorted_coo(nrows, nnz, _p(rows), _p(cols), _p(vals), n, _p(rp), _p(ci), _p(cv))
    return rp, ci, cv


def coo_to_csr(nrows, ncols, rows, cols, vals):
    rows, cols, vals = _i64(rows), _i64(cols), _i64(vals)
    n = rows.size
    rp = np.empty(nrows + 1, np.int64)
    ci = np.empty(n, np.int64)
    cv = np.empty(n, np.int64)
    m = lib().oracle_coo_to_csr(nrows, ncols, n, _p(rows), _p(cols), _p(vals), _p(rp), _p(ci), _p(cv))
    return rp, ci, cv, m


# ---- CSV text (runtime/hydrangea_runtime.h:549-594, runtime/hyd_write_csv.c:5-59) ----
CSV_OK, CSV_PARSE_ERROR, CSV_TOO_FEW = 0, 1, 2


def csv_read(text, count, is_float):
    """The reference's read loop over `text` (bytes): (values, kind, element).
    kind 0 = all `count` values read, 1 = parse error at `element`, 2 = too few
    values (the separator loop hit the end before `element`)."""
    out = np.zeros(max(count, 1), np.float64 if is_float else np.int64)
    err = np.zeros(1, np.int64)
    kind = lib().oracle_csv_read(bytes(text), len(text), count, 1 if is_float else 0,
                                 out.ctypes.data, err.ctypes.data)
    if kind < 0:
        raise RuntimeError("fmemopen failed")
    return out[:count], kind, int(err[0])


def csv_write(vals, rows, row_len, is_float):
    """The reference's write loop: rows x row_len values -> bytes."""
    v = _f64(vals) if is_float else _i64(vals)
    n = np.zeros(1, np.int64)
    p = lib().oracle_csv_write(_p(v), rows, row_len, 1 if is_float else 0, n.ctypes.data)
    if not p:
        raise RuntimeError("open_memstream failed")
    try:
        return ctypes.string_at(p, int(n[0]))
    finally:
        lib().oracle_free(p)


def scan_tokens(strings, is_float):
    """fscanf "%lf" / "%lld" on each string alone: (r, values, consumed)."""
    offs = np.zeros(len(strings) + 1, np.int64)
    offs[1:] = np.cumsum([len(s) for s in strings])
    text = b"".join(strings)
    n = len(strings)
    r = np.zeros(n, np.int32)
    vals = np.zeros(n, np.float64 if is_float else np.int64)
    consumed = np.zeros(n, np.int64)
    lib().oracle_scan_tokens(text, offs.ctypes.data, n, 1 if is_float else 0, r.ctypes.data,
                             vals.ctypes.data, consumed.ctypes.data)
    return r, vals, consumed


def format_values(vals, is_float, stride=40):
    """fprintf "%.17g" / "%lld" text of each value, as a list of bytes."""
    v = _f64(vals) if is_float else _i64(vals)
    n = v.size
    out = np.zeros(n * stride, np.uint8)
    lens = np.zeros(n, np.int32)
    lib().oracle_format_values(_p(v), n, 1 if is_float else 0, _p(out), stride, _p(lens))
    buf = out.tobytes()
    return [buf[k * stride:k * stride + int(lens[k])] for k in range(n)]
