"""Python host side of libhydhip.so (ctypes).

Loads hydrangea-compiler_amd/lib/libhydhip.so (built by build.py) and exposes
  * the kernel-level C ABI (include/hyd_hip.h) over DeviceArray buffers, and
  * the drop-in array ABI (include/hydrangea_runtime.h, hyd_exports.h): the
    hyd_array_t / hyd_tuple_t structs and the hyd_export_* entry points.

There is no CPU fallback: if the library is missing or no GPU is present,
calls raise HydError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhydhip.so")

HYD_OK = 0
HYD_E_INVALID = 1
HYD_E_OOM = 2
HYD_E_HIP = 3
HYD_E_NODEV = 4
HYD_E_REJECTED = 5
HYD_E_TIMEOUT = 6
HYD_E_BUSY = 7

SCATTER_AUTO = 0
SCATTER_LDS = 1
SCATTER_GLOBAL = 2
SCATTER_LDS_SPLIT = 3

STENCIL_HEAT = 0
STENCIL_LAPLACIAN = 1
STENCIL_JACOBI4 = 2
STENCIL_AVG5 = 3

HYD_MAX_DIMS = 8


class HydError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hydhip error %d: %s" % (code, msg))
        self.code = code


# ---- the array ABI (runtime/hydrangea_runtime.h:220-235 of the reference) ----
class HydTuple(ctypes.Structure):
    _fields_ = [("elems", ctypes.c_int64 * HYD_MAX_DIMS), ("ndims", ctypes.c_int)]


class HydStorage(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_ubyte)), ("refcount", ctypes.c_int64)]


class HydArray(ctypes.Structure):
    _fields_ = [
        ("storage", ctypes.POINTER(HydStorage)),
        ("data", ctypes.POINTER(ctypes.c_ubyte)),
        ("elem_size", ctypes.c_size_t),
        ("shape", HydTuple),
    ]


class HydCsrRecordT(ctypes.Structure):
    """hyd_record_col_idx_ai_ncols_i_nnz_i_nrows_i_row_ptr_ai_vals_ai_t"""
    _fields_ = [
        ("col_idx", ctypes.POINTER(HydArray)),
        ("ncols", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("nrows", ctypes.c_int64),
        ("row_ptr", ctypes.POINTER(HydArray)),
        ("vals", ctypes.POINTER(HydArray)),
    ]


class HydRequest(ctypes.Structure):
    """hyd_request_t (include/hyd_serve.h)"""
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("n", ctypes.c_int64),
                ("h", ctypes.c_int64), ("w", ctypes.c_int64), ("iters", ctypes.c_int64),
                ("bins", ctypes.c_int64), ("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64),
                ("in0", ctypes.c_void_p), ("in1", ctypes.c_void_p), ("in2", ctypes.c_void_p),
                ("deadline_ms", ctypes.c_double)]


class HydResult(ctypes.Structure):
    """hyd_result_t"""
    _fields_ = [("status", ctypes.c_int32), ("done", ctypes.c_int32), ("out0", ctypes.c_void_p),
                ("out1", ctypes.c_void_p), ("out2", ctypes.c_void_p), ("count", ctypes.c_int64),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("bytes", ctypes.c_int64), ("batch_id", ctypes.c_int64)]


class CsvStatus(ctypes.Structure):
    _fields_ = [("error_element", ctypes.c_int64), ("error_kind", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("numbers_found", ctypes.c_int64),
                ("exact_path", ctypes.c_int64)]


class HydServerStats(ctypes.Structure):
    """hyd_server_stats_t"""
    _fields_ = [(name, ctypes.c_int64) for name in (
        "budget", "in_use", "peak_in_use", "submitted", "admitted", "completed", "failed", "rejected",
        "timed_out", "queued", "running", "jobs", "batched_jobs", "batched_requests")]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_INT = ctypes.c_int
_D = ctypes.c_double
_PA = ctypes.POINTER(HydArray)

# name -> (restype, argtypes)
_SIGS = {
    "hydhip_last_error": (ctypes.c_char_p, []),
    "hyd_hip_init": (_INT, [ctypes.c_char_p]),
    "hyd_hip_cleanup": (None, []),
    "hydhip_device_info": (_INT, [ctypes.POINTER(_INT), ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                  ctypes.c_char_p, _INT]),
    "hydhip_malloc": (_INT, [ctypes.POINTER(_P), _I64]),
    "hydhip_free": (_INT, [_P]),
    "hydhip_memcpy_h2d": (_INT, [_P, _P, _I64, _P]),
    "hydhip_memcpy_d2h": (_INT, [_P, _P, _I64, _P]),
    "hydhip_memcpy_d2d": (_INT, [_P, _P, _I64, _P]),
    "hydhip_stream_create": (_INT, [ctypes.POINTER(_P)]),
    "hydhip_stream_destroy": (_INT, [_P]),
    "hydhip_stream_sync": (_INT, [_P]),
    "hydhip_device_sync": (_INT, []),
    "hydhip_default_stream": (_P, []),
    "hydhip_event_create": (_INT, [ctypes.POINTER(_P)]),
    "hydhip_event_destroy": (_INT, [_P]),
    "hydhip_event_record": (_INT, [_P, _P]),
    "hydhip_event_elapsed_ms": (_INT, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
    "hydhip_fill_splitmix_i64": (_INT, [_P, _I64, _U64, _U64, _INT, _P]),
    "hydhip_fill_splitmix_f64": (_INT, [_P, _I64, _U64, _U64, _D, _D, _P]),
    "hydhip_weighted_histogram_i64": (_INT, [_P, _I64, _I64, _P, _INT, _P]),
    "hydhip_weighted_histogram_i64_noalloc": (_INT, [_P, _I64, _I64, _P, _INT, _P]),
    "hydhip_scatter_add_i64": (_INT, [_P, _P, _I64, _P, _I64, _INT, _P]),
    "hydhip_scatter_guarded_add_i64": (_INT, [_P, _P, _P, _I64, _P, _I64, _INT, _P]),
    "hydhip_scatter_chain_add_i64": (_INT, [_I64, _P, _P, _P, _P, _P, _I64, _INT, _P]),
    "hydhip_scatter_guarded_add_f64": (_INT, [_P, _P, _P, _I64, _P, _I64, _P]),
    "hydhip_scatter_chain_add_f64": (_INT, [_I64, _P, _P, _P, _P, _P, _I64, _P]),
    "hydhip_dot_f64": (_INT, [_P, _P, _I64, _D, _P, _P]),
    "hydhip_map_chain_f64": (_INT, [_P, _P, _I64, _P, _P]),
    "hydhip_math_f64": (_INT, [_INT, _P, _I64, _P, _P]),
    "hydhip_host_exp": (None, [_P, _I64, _P]),
    "hydhip_host_mod_i64": (_INT, [_P, _I64, _I64, _P]),
    "hydhip_copy_f64": (_INT, [_P, _I64, _P, _P]),
    "hydhip_stencil2d_clamp_f64": (_INT, [_P, _I64, _I64, _INT, _D, _P, _P]),
    "hydhip_stencil2d_iterate_f64": (_INT, [_P, _I64, _I64, _INT, _D, _I64, _P, _P, _INT,
                                            ctypes.POINTER(_P), _P]),
    "hydhip_scan_i64": (_INT, [_P, _I64, _I64, _P, _INT, _INT, _P]),
    "hydhip_sort_indices_i64": (_INT, [_P, _I64, _P, _P]),
    "hydhip_gather_i64": (_INT, [_P, _I64, _P, _I64, _P, _P]),
    "hydhip_segmented_reduce_i64": (_INT, [_P, _I64, _P, _I64, _I64, _P, _P]),
    "hydhip_coo_sum_duplicates_i64": (_INT, [_P, _P, _P, _I64, _I64, _P, _P, _P,
                                             ctypes.POINTER(_I64), _P]),
    "hydhip_csr_from_sorted_coo_i64": (_INT, [_I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "hydhip_coo_to_csr_i64": (_INT, [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P,
                                     ctypes.POINTER(_I64), _P]),
    "hydhip_batched_histogram_i64": (_INT, [_P, _P, _P, _P, _P, _I64, _P, _INT, _P]),
    "hydhip_tune_batched": (_INT, [ctypes.c_char_p, _P]),
    "hydhip_tune_set": (_INT, [_I64, _I64, _INT]),
    "hydhip_tune_clear": (_INT, []),
    "hydhip_tune_table": (_INT, [ctypes.POINTER(_INT), _INT]),
    "hydhip_tune_shapes": (_INT, [ctypes.c_char_p, _P]),
    "hydhip_tune_shape_set": (_INT, [_INT, _I64, _INT]),
    "hydhip_tune_shape_table": (_INT, [ctypes.POINTER(_INT), _INT]),
    "hydhip_batched_dot_f64": (_INT, [_P, _P, _P, _I64, _D, _P, _P]),
    "hydhip_csv_parse_i64": (_INT, [_P, _I64, _I64, _P, ctypes.POINTER(CsvStatus), _P]),
    "hydhip_csv_parse_f64": (_INT, [_P, _I64, _I64, _P, ctypes.POINTER(CsvStatus), _P]),
    "hydhip_csv_text_bound": (_I64, [_I64]),
    "hydhip_csv_format_i64": (_INT, [_P, _I64, _I64, _P, _I64, ctypes.POINTER(_I64), _P]),
    "hydhip_csv_format_f64": (_INT, [_P, _I64, _I64, _P, _I64, ctypes.POINTER(_I64),
                                     ctypes.POINTER(_I64), _P]),
    "hydhip_host_csv_scan": (_INT, [ctypes.c_char_p, _I64, _INT, _INT, _P, ctypes.POINTER(_I64),
                                    ctypes.POINTER(_INT)]),
    "hydhip_host_csv_format": (_I64, [_P, _I64, _INT, _INT, _P, _I64, _P]),
    "hydhip_csv_test_force_exact": (None, [_INT]),
    "hydhip_host_csv_gap_flags": (None, [_P, _I64, _P]),
    "hyd_hip_bench": (_I64, [ctypes.c_char_p, _P, _P, _I64, _I64, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    # serving
    "hyd_server_create": (_P, [_I64, _INT, _INT]),
    "hyd_server_destroy": (None, [_P]),
    "hyd_server_submit": (_INT, [_P, ctypes.POINTER(HydRequest), ctypes.POINTER(_I64)]),
    "hyd_server_poll": (_INT, [_P, _I64, ctypes.POINTER(HydResult)]),
    "hyd_server_wait": (_INT, [_P, _I64, _D, ctypes.POINTER(HydResult)]),
    "hyd_server_wait_all": (_INT, [_P, _D]),
    "hyd_server_next_done": (_INT, [_P, _D, ctypes.POINTER(_I64), ctypes.POINTER(HydResult)]),
    "hyd_server_release": (_INT, [_P, _I64]),
    "hyd_server_stats": (_INT, [_P, ctypes.POINTER(HydServerStats)]),
    "hyd_server_inject_failures": (_INT, [_P, _I64]),
    "hyd_request_bytes": (_I64, [ctypes.POINTER(HydRequest)]),
    "hydhip_run_request_plain": (_INT, [ctypes.POINTER(HydRequest), ctypes.POINTER(HydResult)]),
    "hydhip_result_free_plain": (_INT, [ctypes.POINTER(HydResult)]),
    "hydhip_test_plain_alloc": (_INT, [ctypes.POINTER(HydRequest), _INT, ctypes.POINTER(_I64)]),
    # array ABI
    "hyd_shape_size": (_I64, [HydTuple]),
    "hyd_array_alloc_bytes": (_PA, [HydTuple, ctypes.c_size_t]),
    "hyd_array_alloc": (_PA, [HydTuple]),
    "hyd_array_alloc_copy": (_PA, [_PA]),
    "hyd_array_alloc_device": (_PA, [HydTuple, ctypes.c_size_t]),
    "hyd_array_reshape_view": (_PA, [_PA, HydTuple]),
    "hyd_array_free": (None, [_PA]),
    "hyd_array_length": (_I64, [_PA]),
    "hyd_array_device_data": (_P, [_PA]),
    "hyd_array_sync_host": (_INT, [_PA]),
    "hyd_array_device_data_keep": (_P, [_PA]),
    "hyd_array_mark_device_written": (_INT, [_PA]),
    "hyd_array_enter_host": (_INT, [_PA]),
    "hyd_array_mark_host_dirty": (_INT, [_PA]),
    "hyd_array_is_device_resident": (_INT, [_PA]),
    "hyd_hip_set_host_sync": (None, [_INT]),
    "hyd_hip_get_host_sync": (_INT, []),
    "hyd_hip_set_stream": (None, [_P]),
    "hyd_hip_get_stream": (_P, []),
    "hyd_array_alloc_device8": (_PA, [HydTuple]),
    "hydhip_scratch": (_P, [_INT, _I64, _P]),
    "hydhip_zero_scratch": (_P, [_P]),
    "hyd_sort_indices": (_PA, [_PA]),
    "hyd_read_array_csv": (_PA, [ctypes.c_char_p, HydTuple]),
    "hyd_read_float_array_csv": (_PA, [ctypes.c_char_p, HydTuple]),
    "hyd_write_array_csv": (_I64, [_PA, ctypes.c_char_p]),
    "hyd_write_array_csv_float": (_I64, [_PA, ctypes.c_char_p]),
    "hyd_export_weighted_histogram": (_PA, [_PA]),
    "hyd_export_weighted_histogram_bins": (_PA, [_PA, _I64]),
    "hyd_export_dot": (_D, [_PA, _PA]),
    "hyd_export_map_chain": (_PA, [_PA, _PA]),
    "hyd_export_heat_step": (_PA, [_PA]),
    "hyd_export_jacobi_heat": (_PA, [_PA, _I64]),
    "hyd_export_laplacian": (_PA, [_PA]),
    "hyd_export_coo_to_csr": (HydCsrRecordT, [_I64, _I64, _I64, _PA, _PA, _PA]),
}

_lib = None


def library_path():
    return LIB_PATH


class _TypedLib:
    """Only functions with a declared signature are reachable: ctypes passes an
    undeclared integer argument as a 32-bit int, which would truncate device
    pointers."""

    def __init__(self, cdll):
        self._cdll = cdll

    def __getattr__(self, name):
        if name not in _SIGS:
            raise AttributeError("libhydhip function %s has no declared ctypes signature" % name)
        return getattr(self._cdll, name)


def lib():
    """The loaded libhydhip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HydError(HYD_E_NODEV, "libhydhip.so not built: run hydrangea-compiler_amd/build.py")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = _TypedLib(l)
    return _lib


def _check(rc):
    if rc != HYD_OK:
        raise HydError(rc, lib().hydhip_last_error().decode(errors="replace"))


def init():
    _check(lib().hyd_hip_init(None))


def device_info():
    cus = _INT()
    hbm = _I64()
    lds = _I64()
    name = ctypes.create_string_buffer(128)
    _check(lib().hydhip_device_info(ctypes.byref(cus), ctypes.byref(hbm), ctypes.byref(lds), name, 128))
    return {"num_cus": cus.value, "hbm_bytes": hbm.value, "lds_bytes": lds.value,
            "arch": name.value.decode()}


def sync(stream=None):
    _check(lib().hydhip_stream_sync(stream))


def device_sync():
    _check(lib().hydhip_device_sync())


class DeviceArray:
    """A device (HBM) buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = _P()
        _check(lib().hydhip_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _check(lib().hydhip_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes, stream))
        return d

    def numpy(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(lib().hydhip_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def free(self):
        if self.ptr:
            lib().hydhip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def empty(shape, dtype):
    return DeviceArray(shape, dtype)


# ---- synthetic inputs ---------------------------------------------------------
def splitmix_i64(n, seed, counter0=0, bits=64, stream=None):
    d = DeviceArray((n,), np.int64)
    _check(lib().hydhip_fill_splitmix_i64(d.ptr, n, seed & (2**64 - 1), counter0, 64 - bits, stream))
    return d


def splitmix_f64(n, seed, counter0=0, symmetric=False, stream=None):
    d = DeviceArray((n,), np.float64)
    scale, offset = (2.0, -1.0) if symmetric else (1.0, 0.0)
    _check(lib().hydhip_fill_splitmix_f64(d.ptr, n, seed & (2**64 - 1), counter0, scale, offset, stream))
    return d


# ---- kernels --------------------------------------------------------------------
def weighted_histogram(data, nbins, strategy=SCATTER_AUTO, out=None, stream=None):
    out = out or DeviceArray((nbins,), np.int64)
    _check(lib().hydhip_weighted_histogram_i64(data.ptr, data.size, nbins, out.ptr, strategy, stream))
    return out


def scatter_add(idx, vals, dst, strategy=SCATTER_AUTO, stream=None):
    _check(lib().hydhip_scatter_add_i64(idx.ptr, vals.ptr, idx.size, dst.ptr, dst.size, strategy, stream))
    return dst


def scatter_guarded_add(idx, vals, guard, dst, strategy=SCATTER_AUTO, stream=None):
    """dst (holding the defaults) += vals at idx where guard != 0 (guard None:
    everywhere).  int64 dst: (+); float64 dst: order-exact (+.)."""
    g = guard.ptr if guard is not None else None
    if dst.dtype == np.float64:
        _check(lib().hydhip_scatter_guarded_add_f64(idx.ptr, vals.ptr, g, idx.size, dst.ptr, dst.size,
                                                    stream))
    else:
        _check(lib().hydhip_scatter_guarded_add_i64(idx.ptr, vals.ptr, g, idx.size, dst.ptr, dst.size,
                                                    strategy, stream))
    return dst


def scatter_chain_add(phases, dst, strategy=SCATTER_AUTO, stream=None):
    """phases: list of (idx, vals, guard or None) DeviceArrays, applied in turn."""
    k = len(phases)
    ip = (ctypes.c_void_p * max(k, 1))(*[p[0].ptr for p in phases])
    vp = (ctypes.c_void_p * max(k, 1))(*[p[1].ptr for p in phases])
    gp = (ctypes.c_void_p * max(k, 1))(*[(p[2].ptr if p[2] is not None else None) for p in phases])
    lens = np.array([p[0].size for p in phases] or [0], np.int64)
    if dst.dtype == np.float64:
        _check(lib().hydhip_scatter_chain_add_f64(k, ip, vp, gp, lens.ctypes.data, dst.ptr, dst.size, stream))
    else:
        _check(lib().hydhip_scatter_chain_add_i64(k, ip, vp, gp, lens.ctypes.data, dst.ptr, dst.size,
                                                  strategy, stream))
    return dst


def copy_device(src, dst, stream=None):
    """dst <- src (device to device, same byte size), stream-ordered."""
    if src.nbytes != dst.nbytes:
        raise ValueError("copy_device: sizes differ")
    _check(lib().hydhip_memcpy_d2d(dst.ptr, src.ptr, src.nbytes, stream))
    return dst


def dot(xs, ys, init=0.0, out=None, stream=None):
    """Returns the device scalar (DeviceArray of shape (1,))."""
    out = out or DeviceArray((1,), np.float64)
    _check(lib().hydhip_dot_f64(xs.ptr, ys.ptr, min(xs.size, ys.size), init, out.ptr, stream))
    return out


def map_chain(xs, ys, out=None, stream=None):
    n = min(xs.size, ys.size)
    out = out or DeviceArray((n,), np.float64)
    _check(lib().hydhip_map_chain_f64(xs.ptr, ys.ptr, n, out.ptr, stream))
    return out


def stencil2d(u, body=STENCIL_HEAT, alpha=0.1, out=None, stream=None):
    h, w = u.shape
    out = out or DeviceArray((h, w), np.float64)
    _check(lib().hydhip_stencil2d_clamp_f64(u.ptr, h, w, body, alpha, out.ptr, stream))
    return out


def iterate_stencil2d(u, iters, body=STENCIL_HEAT, alpha=0.1, buf0=None, buf1=None, use_graph=True,
                      stream=None):
    """Returns (result_array, buf0, buf1); result_array is buf0 or buf1."""
    h, w = u.shape
    buf0 = buf0 or DeviceArray((h, w), np.float64)
    buf1 = buf1 or DeviceArray((h, w), np.float64)
    res = _P()
    _check(lib().hydhip_stencil2d_iterate_f64(u.ptr, h, w, body, alpha, iters, buf0.ptr, buf1.ptr,
                                              1 if use_graph else 0, ctypes.byref(res), stream))
    return (buf0 if res.value == buf0.ptr else buf1), buf0, buf1


def scan(a, init=0, inclusive=False, reverse=False, out=None, stream=None):
    out = out or DeviceArray(a.shape, np.int64)
    _check(lib().hydhip_scan_i64(a.ptr, a.size, init, out.ptr, int(inclusive), int(reverse), stream))
    return out


def sort_indices(keys, out=None, stream=None):
    out = out or DeviceArray(keys.shape, np.int64)
    _check(lib().hydhip_sort_indices_i64(keys.ptr, keys.size, out.ptr, stream))
    return out


def gather(perm, src, out=None, stream=None):
    out = out or DeviceArray(perm.shape, np.int64)
    _check(lib().hydhip_gather_i64(perm.ptr, perm.size, src.ptr, src.size, out.ptr, stream))
    return out


def segmented_reduce(offsets, vals, init=0, out=None, stream=None):
    out = out or DeviceArray((max(offsets.size - 1, 0),), np.int64)
    _check(lib().hydhip_segmented_reduce_i64(offsets.ptr, offsets.size, vals.ptr, vals.size, init,
                                             out.ptr, stream))
    return out


def coo_sum_duplicates(rows, cols, vals, nnz, stream=None):
    n = rows.size
    o_r, o_c, o_v = (DeviceArray((n,), np.int64) for _ in range(3))
    cnt = _I64()
    _check(lib().hydhip_coo_sum_duplicates_i64(rows.ptr, cols.ptr, vals.ptr, nnz, n, o_r.ptr, o_c.ptr,
                                               o_v.ptr, ctypes.byref(cnt), stream))
    return o_r, o_c, o_v, cnt.value


def csr_from_sorted_coo(nrows, nnz, rows, cols, vals, stream=None):
    n = cols.size
    rp = DeviceArray((nrows + 1,), np.int64)
    ci = DeviceArray((n,), np.int64)
    cv = DeviceArray((n,), np.int64)
    _check(lib().hydhip_csr_from_sorted_coo_i64(nrows, nnz, rows.ptr, cols.ptr, vals.ptr, n, rp.ptr,
                                                ci.ptr, cv.ptr, stream))
    return rp, ci, cv


def coo_to_csr(nrows, ncols, rows, cols, vals, nnz=None, out=None, stream=None):
    """out: (row_ptr[nrows+1], col_idx[nnz], vals[nnz]) device arrays to
    write into (default: fresh ones)."""
    nnz = rows.size if nnz is None else nnz
    if out is None:
        rp = DeviceArray((nrows + 1,), np.int64)
        ci = DeviceArray((nnz,), np.int64)
        cv = DeviceArray((nnz,), np.int64)
    else:
        rp, ci, cv = out
        if rp.size < nrows + 1 or ci.size < nnz or cv.size < nnz:
            raise ValueError("coo_to_csr: output arrays too small")
    cnt = _I64()
    _check(lib().hydhip_coo_to_csr_i64(nrows, ncols, nnz, rows.ptr, cols.ptr, vals.ptr, rp.ptr, ci.ptr,
                                       cv.ptr, ctypes.byref(cnt), stream))
    return rp, ci, cv, cnt.value


BATCH_TUNED = 0
BATCH_LDS = 1
BATCH_GLOBAL = 2


def batch_offsets(lens, bins):
    """Packed layout of a batch: element offsets and bin offsets."""
    lens = np.ascontiguousarray(lens, np.int64)
    bins = np.ascontiguousarray(bins, np.int64)
    off = np.zeros(lens.size, np.int64)
    hoff = np.zeros(bins.size, np.int64)
    if lens.size > 1:
        off[1:] = np.cumsum(lens)[:-1]
        hoff[1:] = np.cumsum(bins)[:-1]
    return off, hoff


def batched_histogram(data, off, lens, bins, hoff, hist=None, policy=BATCH_TUNED, stream=None):
    """Instance k: hist[hoff[k]:hoff[k]+bins[k]] = histogram of
    data[off[k]:off[k]+lens[k]].  Descriptor arrays are host numpy arrays."""
    off, lens, bins, hoff = (np.ascontiguousarray(a, np.int64) for a in (off, lens, bins, hoff))
    if hist is None:
        hist = DeviceArray((int((hoff + bins).max()) if bins.size else 0,), np.int64)
    _check(lib().hydhip_batched_histogram_i64(data.ptr, off.ctypes.data, lens.ctypes.data,
                                              bins.ctypes.data, hoff.ctypes.data, lens.size, hist.ptr,
                                              policy, stream))
    return hist


def tune_batched(path=None, stream=None):
    rc = lib().hydhip_tune_batched(path.encode() if path else None, stream)
    if rc < 0:
        _check(-rc)
    return rc


def tune_table():
    buf = (_INT * (3 * 256))()
    n = lib().hydhip_tune_table(buf, 256)
    return [(buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]) for i in range(n)]


SHAPE_MAP_CHAIN, SHAPE_COPY = 0, 1
SHAPE_FLAT = 65536  # per_cu value of the flat form (HYD_SHAPE_FLAT)


def tune_shapes(path=None, stream=None):
    """Measure the streaming grid of the map chain and the copy per size
    class; returns the number of classes measured."""
    rc = lib().hydhip_tune_shapes(path.encode() if path else None, stream)
    if rc < 0:
        _check(-rc)
    return rc


def tune_shape_set(kind, n, per_cu):
    _check(lib().hydhip_tune_shape_set(kind, n, per_cu))


def tune_shape_table():
    buf = (_INT * (3 * 64))()
    n = lib().hydhip_tune_shape_table(buf, 64)
    return [(buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]) for i in range(n)]


def batched_dot(xs_list, ys_list, init=0.0, stream=None):
    """Deterministic per-request dots in one launch; returns a DeviceArray of results."""
    count = len(xs_list)
    outs = DeviceArray((max(count, 1),), np.float64)
    xp = (ctypes.c_void_p * count)(*[x.ptr for x in xs_list])
    yp = (ctypes.c_void_p * count)(*[y.ptr for y in ys_list])
    op = (ctypes.c_void_p * count)(*[outs.ptr + 8 * i for i in range(count)])
    lens = np.array([min(x.size, y.size) for x, y in zip(xs_list, ys_list)], np.int64)
    _check(lib().hydhip_batched_dot_f64(xp, yp, lens.ctypes.data, count, init, op, stream))
    return outs


# ---- serving -------------------------------------------------------------------------
REQ_STENCIL = 0
REQ_HIST = 1
REQ_DOT = 2
REQ_COO = 3
SERVE_NO_BATCH = 1
SERVE_FAKE = 2


def request(kind, n=0, h=0, w=0, iters=0, bins=0, nrows=0, ncols=0, in0=None, in1=None, in2=None,
            deadline_ms=0.0):
    return HydRequest(kind, 0, n, h, w, iters, bins, nrows, ncols, in0, in1, in2, deadline_ms)


class Server:
    """Continuous-batching server (include/hyd_serve.h)."""

    def __init__(self, budget_bytes, workers=4, flags=0):
        self.ptr = lib().hyd_server_create(budget_bytes, workers, flags)
        if not self.ptr:
            raise HydError(HYD_E_INVALID, lib().hydhip_last_error().decode(errors="replace"))

    def submit(self, req):
        rid = _I64()
        rc = lib().hyd_server_submit(self.ptr, ctypes.byref(req), ctypes.byref(rid))
        if rc != HYD_OK:
            raise HydError(rc, lib().hydhip_last_error().decode(errors="replace"))
        return rid.value

    def poll(self, rid):
        res = HydResult()
        rc = lib().hyd_server_poll(self.ptr, rid, ctypes.byref(res))
        return (res if rc == HYD_OK else None)

    def wait(self, rid, timeout_ms=60_000.0):
        res = HydResult()
        rc = lib().hyd_server_wait(self.ptr, rid, timeout_ms, ctypes.byref(res))
        if rc != HYD_OK:
            raise HydError(rc, "request %d not finished" % rid)
        return res

    def next_done(self, timeout_ms=0.0):
        """(id, result) of the next finished request, or None on timeout."""
        rid = _I64()
        res = HydResult()
        rc = lib().hyd_server_next_done(self.ptr, timeout_ms, ctypes.byref(rid), ctypes.byref(res))
        return (rid.value, res) if rc == HYD_OK else None

    def wait_all(self, timeout_ms=60_000.0):
        return lib().hyd_server_wait_all(self.ptr, timeout_ms) == HYD_OK

    def release(self, rid):
        _check(lib().hyd_server_release(self.ptr, rid))

    def stats(self):
        st = HydServerStats()
        _check(lib().hyd_server_stats(self.ptr, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in HydServerStats._fields_}

    def inject_failures(self, njobs):
        _check(lib().hyd_server_inject_failures(self.ptr, njobs))

    def close(self):
        if self.ptr:
            lib().hyd_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_bytes(req):
    return lib().hyd_request_bytes(ctypes.byref(req))


def run_plain(req):
    res = HydResult()
    _check(lib().hydhip_run_request_plain(ctypes.byref(req), ctypes.byref(res)))
    return res


def free_plain(res):
    _check(lib().hydhip_result_free_plain(ctypes.byref(res)))


def copy_out(ptr, count, dtype):
    out = np.empty(count, dtype)
    if count:
        _check(lib().hydhip_memcpy_d2h(out.ctypes.data, ptr, out.nbytes, None))
    return out


# ---- events and streams ---------------------------------------------------------
class Event:
    def __init__(self):
        p = _P()
        _check(lib().hydhip_event_create(ctypes.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        _check(lib().hydhip_event_record(self.ptr, stream))

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _check(lib().hydhip_event_elapsed_ms(self.ptr, later.ptr, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().hydhip_event_destroy(self.ptr)
        except Exception:
            pass


def stream_create():
    p = _P()
    _check(lib().hydhip_stream_create(ctypes.byref(p)))
    return p.value


def stream_destroy(s):
    _check(lib().hydhip_stream_destroy(s))


# ---- CSV text <-> arrays ------------------------------------------------------
CSV_OK, CSV_PARSE_ERROR, CSV_TOO_FEW = 0, 1, 2


def csv_parse(text, count, is_float, out=None, stream=None):
    """Parses `text` (a DeviceArray of bytes, or bytes to upload) as the
    reference's read_array / read_array_float loop reads a file.  Returns
    (DeviceArray of `count` values, CsvStatus)."""
    if not isinstance(text, DeviceArray):
        text = DeviceArray.from_numpy(np.frombuffer(bytes(text), np.uint8))
    dtype = np.float64 if is_float else np.int64
    out = out if out is not None else DeviceArray((count,), dtype)
    st = CsvStatus()
    fn = lib().hydhip_csv_parse_f64 if is_float else lib().hydhip_csv_parse_i64
    _check(fn(text.ptr, text.nbytes, count, out.ptr, ctypes.byref(st), stream))
    return out, st


def csv_format(vals, row_len=None, stream=None, with_exact=False):
    """Text the reference's write_array / write_array_float would write for
    the values of `vals` (DeviceArray, int64 or float64), rows of `row_len`."""
    n = vals.size
    row_len = n if row_len is None else row_len
    cap = max(1, lib().hydhip_csv_text_bound(n))
    text = DeviceArray((cap,), np.uint8)
    nb = _I64()
    exact = _I64()
    if vals.dtype == np.float64:
        _check(lib().hydhip_csv_format_f64(vals.ptr, n, row_len, text.ptr, cap, ctypes.byref(nb),
                                           ctypes.byref(exact), stream))
    else:
        _check(lib().hydhip_csv_format_i64(vals.ptr, n, row_len, text.ptr, cap, ctypes.byref(nb), stream))
    out = copy_out(text.ptr, nb.value, np.uint8).tobytes() if nb.value else b""
    text.free()
    return (out, exact.value) if with_exact else out


def host_csv_scan(strings, is_float, force_exact=False, swar=False):
    """CPU build of the device scanner on each string: (r, values, consumed, exact)."""
    n = len(strings)
    r = np.zeros(n, np.int32)
    vals = np.zeros(n, np.float64 if is_float else np.int64)
    consumed = np.zeros(n, np.int64)
    exact = np.zeros(n, np.int32)
    c = _I64()
    e = _INT()
    for k, s in enumerate(strings):
        r[k] = lib().hydhip_host_csv_scan(s, len(s), 1 if is_float else 0,
                                          (1 if force_exact else 0) | (2 if swar else 0),
                                          vals.ctypes.data + 8 * k, ctypes.byref(c), ctypes.byref(e))
        consumed[k] = c.value
        exact[k] = e.value
    return r, vals, consumed, exact


def host_csv_format(vals, is_float, stride=40, force_exact=False):
    """CPU build of the device formatter: (list of bytes, exact-path count)."""
    v = np.ascontiguousarray(vals, np.float64 if is_float else np.int64)
    n = v.size
    out = np.zeros(max(1, n * stride), np.uint8)
    lens = np.zeros(max(1, n), np.int32)
    exact = lib().hydhip_host_csv_format(v.ctypes.data, n, 1 if is_float else 0, 1 if force_exact else 0,
                                         out.ctypes.data, stride, lens.ctypes.data)
    buf = out.tobytes()
    return [buf[k * stride:k * stride + int(lens[k])] for k in range(n)], exact


def default_stream():
    return lib().hydhip_default_stream()


# ---- drop-in array ABI helpers ----------------------------------------------------
def hyd_tuple(*dims):
    t = HydTuple()
    t.ndims = len(dims)
    for i, d in enumerate(dims):
        t.elems[i] = int(d)
    return t


def host_array(a):
    """A hyd_array_t* (allocated by the library) holding a copy of numpy `a`."""
    a = np.ascontiguousarray(a)
    arr = lib().hyd_array_alloc_bytes(hyd_tuple(*a.shape), a.dtype.itemsize)
    if a.nbytes:
        ctypes.memmove(arr.contents.data, a.ctypes.data, a.nbytes)
    return arr


def read_array(arr, dtype):
    """numpy copy of a hyd_array_t*'s host bytes."""
    a = arr.contents
    shape = tuple(a.shape.elems[i] for i in range(a.shape.ndims))
    n = int(np.prod(shape, dtype=np.int64))
    out = np.empty(shape, dtype)
    if n:
        ctypes.memmove(out.ctypes.data, a.data, n * np.dtype(dtype).itemsize)
    return out
