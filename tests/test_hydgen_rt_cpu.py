"""The host side of include/hydgen_rt.h (CPU): the multiply-shift division
the generated kernels use for divisors that do not change over a loop equals
C's / and % (truncation toward zero) on edge values and random operands, and
the shape-tuple helpers equal the runtime's hyd_nd_to_flat / hyd_flat_to_nd /
hyd_shape_size.  Built with hipcc as host code; no GPU call is made."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "hydgen_rt.h"
#include <random>
int main() {
    const int64_t edges[] = {0, 1, -1, 2, -2, 3, -3, 7, -7, 63, 64, 65, -64, 1000003, -1000003,
                             (int64_t)1 << 31, -((int64_t)1 << 31), ((int64_t)1 << 62) + 1, INT64_MAX, INT64_MAX - 1,
                             INT64_MIN + 1, INT64_MIN};
    std::mt19937_64 rng(7);
    long bad = 0, checked = 0;
    auto check = [&](int64_t x, int64_t m) {
        if (m == 0 || (x == INT64_MIN && m == -1)) return;  /* undefined in C */
        const hydk_divc_t d = hydk_make_divc(m);
        if (hydk_cdiv(x, d) != x / m || hydk_cmod(x, d) != x % m) {
            if (bad < 5) printf("mismatch x=%lld m=%lld\n", (long long)x, (long long)m);
            ++bad;
        }
        ++checked;
    };
    for (int64_t x : edges) for (int64_t m : edges) check(x, m);
    for (int k = 0; k < 2000000; ++k) {
        const int64_t x = (int64_t)rng();
        int64_t m = (int64_t)(rng() >> (rng() % 64));
        if (rng() & 1) m = -m;
        check(x, m);
        check(x >> (rng() % 64), m);
    }
    /* tuples against the runtime's own helpers */
    for (int k = 0; k < 200000; ++k) {
        const int nd = 1 + (int)(rng() % 8);
        hyd_tuple_t s = hydk_tuple(nd), t = hydk_tuple(nd);
        for (int i = 0; i < nd; ++i) {
            s.elems[i] = 1 + (int64_t)(rng() % 9);
            t.elems[i] = (int64_t)(rng() % s.elems[i]);
        }
        const int64_t f = hyd_nd_to_flat(t, s);
        if (hydk_nd_to_flat(t, s) != f || hydk_shape_size(s) != hyd_shape_size(s) ||
            hydk_shape_last(s) != hyd_shape_last(s)) ++bad;
        const hyd_tuple_t back = hydk_flat_to_nd(f, s), ref = hyd_flat_to_nd(f, s);
        for (int i = 0; i < 8; ++i) if (back.elems[i] != ref.elems[i]) ++bad;
        const hyd_tuple_t init = hydk_shape_init(s), rinit = hyd_shape_init(s);
        if (init.ndims != rinit.ndims) ++bad;
        for (int i = 0; i < 8; ++i) if (init.elems[i] != rinit.elems[i]) ++bad;
        ++checked;
    }
    printf("checked=%ld bad=%ld\n", checked, bad);
    return bad != 0;
}
'''


def test_divisor_constants_and_tuple_helpers(tmp_path):
    src = tmp_path / "rt.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "rt"
    lib = os.path.join(ROOT, "hydrangea-compiler_amd", "lib")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L" + lib, "-lhydhip", "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "hipcc" in r.stderr and "not found" in r.stderr:
        pytest.skip("hipcc missing")
    assert r.returncode == 0, r.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "bad=0" in run.stdout
