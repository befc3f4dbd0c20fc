"""CPU tests of hydhip-cc, the compile driver the HIP backend runs
(HIPBackend.compileAndRunHIP passes it the printed C and the offload sites):
it builds a program that calls generated launchers and keeps OpenMP for the
host loops, stops with hydgen's message on a bad site, and the program it
builds fails loudly when no GPU is present."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "hydrangea-compiler_amd")
DRIVER = os.path.join(PKG, "hydhip-cc")
CFG = os.path.join(HERE, "golden", "cfg")


def sites_file(tmp_path, *names):
    p = tmp_path / "prog.sites"
    p.write_text("".join(open(os.path.join(CFG, n + ".sites")).read() for n in names))
    return str(p)


def hipbackend_args(cc, parallel, cfile, sites, exe):
    """HIPBackend.hydhipCcArgs, restated."""
    return [cfile, sites, "-o", exe] + (["--cc", cc] if cc else []) + ([] if parallel else ["--no-openmp"])


def run_driver(args, **env):
    e = dict(os.environ, HYD_HOST_MARCH="x86-64-v3", **env)
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, env=e)


def test_builds_program_with_openmp_and_launchers(tmp_path):
    exe = str(tmp_path / "prog")
    r = run_driver(hipbackend_args("gcc", True, os.path.join(CFG, "dotmap_program.c"),
                                   sites_file(tmp_path, "map_chain", "dot"), exe))
    assert r.returncode == 0, r.stderr
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libhydhip.so" in ldd and "libamdhip64" in ldd and "libgomp" in ldd
    nm = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    assert "hydk_map_chain_0" in nm and "hydk_dot_0" in nm


def test_keep_writes_generated_sources(tmp_path):
    keep = tmp_path / "keep"
    r = run_driver(hipbackend_args(None, True, os.path.join(CFG, "dotmap_program.c"),
                                   sites_file(tmp_path, "map_chain", "dot"), str(tmp_path / "p")) +
                   ["--keep", str(keep)], CC="gcc")
    assert r.returncode == 0, r.stderr
    assert "hk_kernel_hydk_map_chain_0" in (keep / "kernels.hip").read_text()
    assert "double hydk_dot_0(double v_acc, int64_t v_red_dim" in (keep / "hydk_sites.h").read_text()


def test_bad_site_stops_the_build(tmp_path):
    bad = tmp_path / "bad.sites"
    bad.write_text("HipSite {hsName = \"x\"\n")
    r = run_driver(hipbackend_args("gcc", True, os.path.join(CFG, "dotmap_program.c"), str(bad),
                                   str(tmp_path / "p")))
    assert r.returncode == 1 and "hydgen:" in r.stderr
    assert not (tmp_path / "p").exists()


def test_usage_errors():
    r = run_driver(["only_one_arg.c"])
    assert r.returncode == 2 and "usage" in r.stderr


def test_program_fails_loudly_without_gpu():
    exe = os.path.join(PKG, "_build", "e2e", "dotmap_program")
    if not os.path.exists(exe):
        pytest.skip("build first")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present here")
    r = subprocess.run([exe, "10"], capture_output=True, text=True)
    assert r.returncode == 1 and "no HIP device" in r.stderr
