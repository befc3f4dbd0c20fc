"""Host helpers of the array runtime that emitted programs call, from C, on the
CPU (no GPU): printing (runtime/hydrangea_runtime.h:458-548), slices
(:277-290), element access, get_env_int / get_env_string
(:601-624, runtime/hyd_write_csv.c:61-72).  Expected text is the format
those functions print, read from the reference header."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hydrangea-compiler_amd", "lib")

SRC = r"""
#include <stdio.h>
#include <string.h>
#include "hydrangea_runtime.h"

int main(int argc, char** argv) {
    if (argc > 1 && strcmp(argv[1], "env") == 0) {
        printf("%lld %s\n", (long long)getenv_int("HYD_T_INT"), getenv_string("HYD_T_STR"));
        printf("%lld %s\n", (long long)hyd_get_env_int("HYD_T_INT"), hyd_get_env_string("HYD_T_STR"));
        return 0;
    }
    if (argc > 1 && strcmp(argv[1], "unset") == 0) {
        getenv_int("HYD_T_NOT_SET_ANYWHERE");
        return 0;
    }
    hyd_array_t* a = hyd_array_alloc(hyd_tuple_make(2, (int64_t)2, (int64_t)3));
    for (int i = 0; i < 6; i++) hyd_array_data_int64(a)[i] = (int64_t)(i * i) - 3;
    hyd_print_array(a);
    hyd_array_t* f = hyd_array_alloc(hyd_tuple_make(1, (int64_t)4));
    hyd_array_set_float(f, 0, 0.1);
    hyd_array_set_float(f, 1, -2.5);
    hyd_array_set_float(f, 2, 1e300);
    hyd_array_set_float(f, 3, hyd_array_get_float(f, 1) * 2.0);
    hyd_print_float_array(f);
    hyd_array_t* e = hyd_array_alloc(hyd_tuple_make(1, (int64_t)0));
    hyd_print_array(e);
    hyd_print_tuple(hyd_tuple_make(3, (int64_t)7, (int64_t)-1, (int64_t)0));
    hyd_print_int(-42);
    hyd_print_double(1.0 / 3.0);
    hyd_print_pair((int64_t)5, 0.25);
    hyd_print_pair(1.5, (int64_t)-7);
    hyd_tuple_t s = hyd_slice_shape(hyd_tuple_make(3, (int64_t)4, (int64_t)5, (int64_t)6),
                                    hyd_tuple_make(6, (int64_t)-1, (int64_t)-1, (int64_t)1, (int64_t)3,
                                                   (int64_t)0, (int64_t)6));
    hyd_print_tuple(s);
    printf("[");
    hyd_print_int_inline(3);
    printf(" ");
    hyd_print_double_inline(2.0);
    printf(" ");
    hyd_print_array_inline(e);
    printf(" ");
    hyd_print_float_array_inline(f);
    printf("]\n");
    hyd_array_free(a);
    hyd_array_free(f);
    hyd_array_free(e);
    return 0;
}
"""

EXPECTED = """[-3, -2, 1, 6, 13, 22] (shape: [2, 3])
[0.10000000000000001, -2.5, 1.0000000000000001e+300, -5] (shape: [4])
[] (shape: [0])
(7, -1, 0)
-42
0.33333333333333331
(5, 0.25)
(1.5, -7)
(4, 3, 6)
[3 2 [] (shape: [0]) [0.10000000000000001, -2.5, 1.0000000000000001e+300, -5] (shape: [4])]
"""


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("rt")
    (d / "rt.c").write_text(SRC)
    out = str(d / "rt")
    subprocess.run(["gcc", "-std=c11", "-O1", "-I" + INCLUDE, str(d / "rt.c"), "-o", out, "-L" + LIBDIR,
                    "-lhydhip", "-Wl,-rpath," + LIBDIR], check=True)
    return out


def test_print_and_shape_helpers(exe):
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert p.stdout == EXPECTED


def test_env_helpers(exe):
    env = dict(os.environ, HYD_T_INT="123", HYD_T_STR="abc")
    p = subprocess.run([exe, "env"], capture_output=True, text=True, env=env, timeout=60)
    assert p.returncode == 0, p.stderr
    assert p.stdout == "123 abc\n123 abc\n"


def test_env_unset_exits_like_reference(exe):
    p = subprocess.run([exe, "unset"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert "environment variable HYD_T_NOT_SET_ANYWHERE not set" in p.stderr
