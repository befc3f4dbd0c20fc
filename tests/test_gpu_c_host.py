"""The drop-in boundary from C, as the reference's compiled programs use it:
a host built with gcc against include/ and -lhydhip, written the way the C
backend emits bench/stencil/stencil_2d.hyd (get_env_int, read_array_float
[h, w] from a CSV file, the clamped Laplacian, write_array_float), checked
byte for byte against the oracle's read -> stencil -> write; and the
reference's error exits (runtime/hydrangea_runtime.h:549-594: message on
stderr, exit status 1)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hydrangea-compiler_amd", "lib")

HOST_C = r"""
#include <stdio.h>
#include "hydrangea_runtime.h"
#include "hyd_exports.h"

static void laplacian_once(void* arg) {
    hyd_array_t* r = hyd_export_laplacian((hyd_array_t*)arg);
    hyd_array_free(r);
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    if (argv[3][0] == 'b') {  /* the --bench scaffold around the Laplacian */
        hyd_array_t* input = hyd_read_float_array_csv(argv[1], hyd_tuple_make(2, (int64_t)64, (int64_t)64));
        double mn = -1.0, mean = -1.0;
        hyd_hip_bench("laplacian", laplacian_once, input, 2, 5, &mn, &mean);
        printf("%s\n", (mn > 0.0 && mean >= mn) ? "ok" : "bad");
        hyd_array_free(input);
        return 0;
    }
    if (argv[3][0] == 'i') {  /* read_array [n] and write it back */
        int64_t n = hyd_get_env_int("CSV_N");
        hyd_array_t* a = hyd_read_array_csv(argv[1], hyd_tuple_make(1, n));
        hyd_write_array_csv(a, argv[2]);
        hyd_array_free(a);
        return 0;
    }
    int64_t h = hyd_get_env_int("STENCIL_H");
    int64_t w = hyd_get_env_int("STENCIL_W");
    hyd_array_t* input = hyd_read_float_array_csv(argv[1], hyd_tuple_make(2, h, w));
    hyd_array_t* laplacian = hyd_export_laplacian(input);
    hyd_write_array_csv_float(laplacian, argv[2]);
    hyd_array_free(laplacian);
    hyd_array_free(input);
    hyd_hip_cleanup();
    return 0;
}
"""


@pytest.fixture(scope="module")
def host_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("chost")
    src = d / "stencil_2d_host.c"
    src.write_text(HOST_C)
    exe = str(d / "stencil_2d_host")
    subprocess.run(["gcc", "-std=c99", "-O2", "-I" + INCLUDE, str(src), "-o", exe, "-L" + LIBDIR, "-lhydhip",
                    "-Wl,-rpath," + LIBDIR], check=True)
    return exe


def run(exe, args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe] + args, capture_output=True, text=True, env=e, timeout=120)


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (257, 300), (1024, 1536)])
def test_stencil_2d_program_from_c(orc, host_exe, tmp_path, h, w):
    print("stencil_2d host h=%d w=%d" % (h, w), flush=True)
    rng = np.random.default_rng(h * 7919 + w)
    u = rng.standard_normal((h, w)) * 10.0 ** rng.integers(-3, 4, (h, w))
    inp, out = tmp_path / "input.csv", tmp_path / "out.csv"
    inp.write_bytes(orc.csv_write(u.ravel(), h, w, True))
    p = run(host_exe, [str(inp), str(out), "f"], {"STENCIL_H": str(h), "STENCIL_W": str(w)})
    assert p.returncode == 0, p.stderr
    vals, kind, _ = orc.csv_read(inp.read_bytes(), h * w, True)
    assert kind == 0
    lap = orc.stencil2d(vals.reshape(h, w), body=1)
    assert out.read_bytes() == orc.csv_write(lap.ravel(), h, w, True)


def test_read_array_int_round_trip_from_c(orc, host_exe, tmp_path):
    i = np.random.default_rng(3).integers(-2**63, 2**63 - 1, 100_000, dtype=np.int64)
    inp, out = tmp_path / "in.csv", tmp_path / "out.csv"
    inp.write_bytes(b"  " + b" ,\n".join(str(x).encode() for x in i) + b"\r\n")
    p = run(host_exe, [str(inp), str(out), "i"], {"CSV_N": str(i.size)})
    assert p.returncode == 0, p.stderr
    assert out.read_bytes() == orc.csv_write(i, 1, i.size, False)


@pytest.mark.parametrize("text,n,message", [
    (b"1,two,3", 3, "read_array: parse error at element 1"),
    (b"1,2,", 3, "read_array: too few values"),
    (b"1\t,2", 2, "read_array: parse error at element 1"),
])
def test_reference_error_exits_from_c(host_exe, tmp_path, text, n, message):
    inp = tmp_path / "bad.csv"
    inp.write_bytes(text)
    p = run(host_exe, [str(inp), str(tmp_path / "o.csv"), "i"], {"CSV_N": str(n)})
    assert p.returncode == 1
    assert message in p.stderr


def test_missing_file_exits_from_c(host_exe, tmp_path):
    missing = str(tmp_path / "nope.csv")
    p = run(host_exe, [missing, str(tmp_path / "o.csv"), "f"], {"STENCIL_H": "2", "STENCIL_W": "2"})
    assert p.returncode == 1
    assert "read_array_float: cannot open '%s'" % missing in p.stderr


def test_bench_scaffold_from_c(orc, host_exe, tmp_path):
    # CodegenC.hs:515-580: "benchmark[<name>]: min=%.3f ms  mean=%.3f ms" on stderr
    import re
    inp = tmp_path / "input.csv"
    inp.write_bytes(orc.csv_write(np.arange(64 * 64, dtype=np.float64), 64, 64, True))
    p = run(host_exe, [str(inp), str(tmp_path / "o.csv"), "b"])
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip() == "ok"
    assert re.search(r"^benchmark\[laplacian\]: min=\d+\.\d{3} ms  mean=\d+\.\d{3} ms$", p.stderr, re.M), p.stderr
