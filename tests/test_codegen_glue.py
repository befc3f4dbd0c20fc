"""The HIP offload glue's consistency on the CPU: every helper that the
offload pass (hydrangea-compiler_amd/haskell/Language/Hydrangea/CodegenHIP.hs)
emits is defined in include/hyd_hip_glue.h, the header compiles as C99 with
warnings as errors, every library function it calls is exported by
libhydhip.so, and the backend driver links where the library is built.  The
Haskell modules themselves are not compiled (no GHC); the helpers run on the
GPU in test_gpu_glue.py."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hydrangea-compiler_amd", "lib")
HS = os.path.join(ROOT, "hydrangea-compiler_amd", "haskell")
GLUE = os.path.join(INCLUDE, "hyd_hip_glue.h")


def _read(p):
    with open(p) as f:
        return f.read()


def test_emitted_helpers_are_defined():
    hs = _read(os.path.join(HS, "Language", "Hydrangea", "CodegenHIP.hs"))
    emitted = set(re.findall(r'"(hydg_\w+)"', hs))
    defined = set(re.findall(r"static inline \w+\*? (hydg_\w+)\(", _read(GLUE)))
    assert emitted == {"hydg_scatter_add_i64", "hydg_scatter_add_f64", "hydg_weighted_histogram_i64", "hydg_dot_f64"}
    assert emitted <= defined


def test_driver_links_where_the_library_is_built():
    hs = _read(os.path.join(HS, "HIPBackend.hs"))
    assert '"hydrangea-compiler_amd" </> "lib"' in hs and '"-lhydhip"' in hs


def test_glue_header_compiles(tmp_path):
    src = tmp_path / "g.c"
    src.write_text('#include "hyd_hip_glue.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I" + INCLUDE, "-c", str(src), "-o",
                    str(tmp_path / "g.o")], check=True)


def test_glue_calls_are_exported():
    so = os.path.join(LIBDIR, "libhydhip.so")
    if not os.path.exists(so):
        pytest.skip("libhydhip.so not built")
    body = _read(GLUE)
    called = set(re.findall(r"\b(hydhip_\w+|hyd_array_\w+|hyd_shape_size)\(", body))
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert called and called <= exported, sorted(called - exported)
