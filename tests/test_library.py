"""CPU checks of the C-ABI library: it loads, exports every function the
public headers declare, keeps the reference's hyd_array_t layout, and runs
its pure host functions.  No GPU work is launched here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import hydhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in sorted(os.listdir(INCLUDE)):
        # headers of static inline helpers, not library symbols
        if not f.endswith(".h") or f in ("hyd_exp.h", "hydgen_rt.h"):
            continue
        text = open(os.path.join(INCLUDE, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"^\s*#define(?:[^\n]*\\\n)*[^\n]*", "", text, flags=re.M)  # macros, with continuations
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "sizeof", "return"):
                continue
            names.add(name)
    return sorted(names)


def test_library_loads():
    assert os.path.exists(hydhip.LIB_PATH), "build with hydrangea-compiler_amd/build.py"
    hydhip.lib()


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 40
    lib = ctypes.CDLL(hydhip.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_symbol_has_a_ctypes_signature():
    missing = [n for n in declared_functions() if n.startswith(("hydhip_", "hyd_export_", "hyd_hip_"))
               and n not in hydhip._SIGS]
    assert not missing, missing


def test_untyped_calls_are_refused():
    # an undeclared ctypes call would pass device pointers as 32-bit ints
    with pytest.raises(AttributeError):
        hydhip.lib().hyd_tuple_make


def test_struct_layout_matches_reference():
    # runtime/hydrangea_runtime.h:220-235: int64 elems[8] + int ndims (padded),
    # storage {data*, refcount}, array {storage*, data*, elem_size, shape}
    assert ctypes.sizeof(hydhip.HydTuple) == 72
    assert ctypes.sizeof(hydhip.HydArray) == 8 + 8 + 8 + 72
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "hydrangea_runtime.h"
int main(void){printf("%zu %zu %zu %zu\n", sizeof(hyd_tuple_t), sizeof(hyd_array_t),
 offsetof(hyd_array_t, shape), sizeof(hyd_array_storage_t)); return 0;}
'''
    tmp = "/tmp/hyd_layout_check"
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-std=c11", "-I" + INCLUDE, tmp + ".c", "-o", tmp], check=True)
    out = subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["72", "96", "24", "16"]


def test_host_runtime_functions_without_gpu():
    lib = hydhip.lib()
    t = hydhip.hyd_tuple(3, 4)
    assert lib.hyd_shape_size(t) == 12
    arr = lib.hyd_array_alloc_bytes(t, 8)
    a = arr.contents
    assert a.elem_size == 8 and a.shape.ndims == 2 and a.storage.contents.refcount == 1
    # calloc semantics: zero-filled
    assert not hydhip.read_array(arr, np.int64).any()
    v = lib.hyd_array_reshape_view(arr, hydhip.hyd_tuple(12))
    assert a.storage.contents.refcount == 2
    lib.hyd_array_free(v)
    assert a.storage.contents.refcount == 1
    lib.hyd_array_free(arr)


def test_reshape_mismatch_exits_like_reference():
    code = (
        "import sys; sys.path.insert(0, %r); import hydhip; l = hydhip.lib();"
        "a = l.hyd_array_alloc_bytes(hydhip.hyd_tuple(3), 8);"
        "l.hyd_array_reshape_view(a, hydhip.hyd_tuple(4))" % os.path.dirname(hydhip.__file__)
    )
    p = subprocess.run(["python3", "-c", code], capture_output=True, text=True)
    assert p.returncode == 1
    assert "reshape view size mismatch (3 elements to 4)" in p.stderr


def test_host_exp_matches_libm(orc):
    # The map chain's exp (include/hyd_exp.h) compiled for the host must equal the
    # libm exp the emitted C calls, bit for bit, over the whole range.
    rng = np.random.default_rng(7)
    x = np.concatenate([
        rng.uniform(-760.0, 760.0, 1_000_000),
        rng.uniform(0.0, 0.003, 1_000_000),        # the map chain's argument range
        rng.uniform(-1.0, 1.0, 500_000),
        np.frombuffer(rng.integers(0, 2**63, 500_000, dtype=np.uint64).tobytes(), np.float64),
        np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 709.78, 709.79, -745.13, -745.14, -708.4,
                  2.0**-54, 2.0**-55, 512.0, -512.0, 1023.9, -1074.0]),
    ])
    lib = hydhip.lib()
    got = np.empty_like(x)
    lib.hydhip_host_exp(x.ctypes.data, x.size, got.ctypes.data)
    ref = orc.libm_exp(x)
    same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), x[~same][:5]


def c_rem(x, m):
    """C's truncating remainder (sign of the dividend), numpy int64."""
    r = np.abs(x) % m  # |INT64_MIN| wraps to itself; handled below
    r = np.where(x == np.iinfo(np.int64).min, (-(x + m)) % m, r)
    return np.where(x < 0, -r, r)


def test_host_mod_matches_c_remainder():
    # The histograms route with a division-free remainder (multiply by a
    # precomputed magic number); it must equal C's `%` for every int64 and
    # every divisor the kernels accept (1 <= m <= 2^31).
    rng = np.random.default_rng(11)
    lo, hi = np.iinfo(np.int64).min, np.iinfo(np.int64).max
    edge = np.array([0, 1, -1, 2, -2, 63, 64, 65, -63, -64, -65, lo, lo + 1, hi, hi - 1,
                     2**31 - 1, 2**31, -2**31, -2**31 - 1, 2**32, 2**32 + 1, -2**32,
                     2**53 + 1, -(2**53) - 1, 2**62, -(2**62)], np.int64)
    xs = np.concatenate([edge, rng.integers(lo, hi, 200_000, dtype=np.int64, endpoint=True),
                         rng.integers(-2**20, 2**20, 50_000, dtype=np.int64)])
    divisors = [1, 2, 3, 5, 7, 10, 63, 64, 65, 100, 255, 256, 1000, 4095, 4096, 4097, 65535,
                65536, 65537, 1 << 20, (1 << 20) + 7, 2**31 - 1, 2**31]
    divisors += [int(v) for v in rng.integers(2, 2**31, 20)]
    lib = hydhip.lib()
    got = np.empty_like(xs)
    for m in divisors:
        assert lib.hydhip_host_mod_i64(xs.ctypes.data, xs.size, m, got.ctypes.data) == hydhip.HYD_OK
        ref = c_rem(xs, m)
        bad = got != ref
        assert not bad.any(), (m, xs[bad][:4], got[bad][:4], ref[bad][:4])
    # exact Python big-integer check of the edge values
    for m in (3, 64, 2**31 - 1, 2**31):
        lib.hydhip_host_mod_i64(edge.ctypes.data, edge.size, m, got.ctypes.data)
        for x, g in zip(edge.tolist(), got[:edge.size].tolist()):
            r = abs(x) % m
            assert g == (-r if x < 0 else r), (x, m, g)


def test_no_gpu_reports_nodev():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    rc = hydhip.lib().hyd_hip_init(None)
    assert rc == hydhip.HYD_E_NODEV
    with pytest.raises(hydhip.HydError):
        hydhip.DeviceArray((4,), np.int64)
