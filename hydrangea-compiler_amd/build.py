"""Build libhydhip.so for gfx950 with hipcc (in-tree, no cmake).

    python hydrangea-compiler_amd/build.py

compiles every csrc/*.hip to an object in _build/ and links
lib/libhydhip.so.  Every HIP source is compiled with
  -ffp-contract=off      (no a*b+c contraction: the emitted C's ISO semantics)
  -Xclang -target-feature -Xclang -packed-fp32-ops   (no packed fp32 ops)
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libhydhip.so")

ARCH = os.environ.get("HYD_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(HERE, "..", "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libhydhip.so)")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is noise
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


HYDGEN_SRC = os.path.join(HERE, "hydgen")
HYDGEN = os.path.join(HERE, "bin", "hydgen")
PROGRAM_SITES = os.path.join(HERE, "..", "tests", "golden", "cfg")
PROGRAMS_LIB = os.path.join(LIBDIR, "libhydgen_programs.so")
PROGRAMS_SIG = os.path.join(LIBDIR, "libhydgen_programs.sig")


def build_hydgen(verbose=False):
    """The kernel generator (host C++): bin/hydgen."""
    os.makedirs(os.path.dirname(HYDGEN), exist_ok=True)
    srcs = sorted(os.path.join(HYDGEN_SRC, f) for f in os.listdir(HYDGEN_SRC) if f.endswith(".cpp"))
    tmp = HYDGEN + ".tmp"
    out = _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", tmp] + srcs)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, HYDGEN)
    return HYDGEN


def build_programs(verbose=False, jobs=None):
    """The HIP backend's output for the example programs: every
    tests/golden/cfg/*.sites through hydgen, compiled for gfx950 and linked
    into lib/libhydgen_programs.so (with the launchers' signatures beside it)."""
    gen = os.path.join(BUILD, "gen")
    os.makedirs(gen, exist_ok=True)
    cc = hipcc()
    files = sorted(f for f in os.listdir(PROGRAM_SITES) if f.endswith(".sites"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    sigs = []

    def one(f):
        base = os.path.join(gen, f[:-6])
        _run([HYDGEN, os.path.join(PROGRAM_SITES, f), "-o", base + ".hip", "--signatures", base + ".sig"])
        _run([cc] + HIPCC_FLAGS + ["-c", base + ".hip", "-o", base + ".o"])
        return base

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        bases = list(ex.map(one, files))
    for b in bases:
        with open(b + ".sig") as fh:
            sigs.append(fh.read())
    tmp = PROGRAMS_LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [b + ".o" for b in bases] +
         ["-L" + LIBDIR, "-lhydhip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, PROGRAMS_LIB)
    with open(PROGRAMS_SIG, "w") as fh:
        fh.write("".join(sigs))
    if verbose:
        print("built", PROGRAMS_LIB, "(%d programs)" % len(bases))
    return PROGRAMS_LIB


HYDHIP_CC = os.path.join(HERE, "hydhip-cc")
E2E_DIR = os.path.join(BUILD, "e2e")
E2E_PROGRAM = os.path.join(E2E_DIR, "dotmap_program")


def build_e2e(verbose=False):
    """An end-to-end program: tests/golden/cfg/dotmap_program.c (C in the shapes
    the C backend prints) with the map-chain and dot sites, built by hydhip-cc
    as the HIP backend builds a program (host code for x86-64-v3, so that it
    runs on any current x86-64 host)."""
    os.makedirs(E2E_DIR, exist_ok=True)
    sites = os.path.join(E2E_DIR, "dotmap.sites")
    with open(sites, "w") as out:
        for f in ("map_chain.sites", "dot.sites"):
            with open(os.path.join(PROGRAM_SITES, f)) as fh:
                out.write(fh.read())
    env = dict(os.environ, HYD_HOST_MARCH="x86-64-v3")
    p = subprocess.run([HYDHIP_CC, os.path.join(PROGRAM_SITES, "dotmap_program.c"), sites, "-o", E2E_PROGRAM,
                        "--cc", "gcc"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
    if p.returncode != 0:
        raise RuntimeError("hydhip-cc failed:\n" + p.stdout)
    if verbose:
        print("built", E2E_PROGRAM)
    return E2E_PROGRAM


def build(verbose=False, jobs=None):
    """Compile every HIP source and link lib/libhydhip.so.  Returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cc = hipcc()
    srcs = sources()
    objs = [os.path.join(BUILD, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)

    def compile_one(pair):
        src, obj = pair
        cmd = [cc] + HIPCC_FLAGS + ["-c", src, "-o", obj]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_hydgen(verbose)
    build_programs(verbose, jobs)
    build_e2e(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
