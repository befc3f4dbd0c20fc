"""Build the CPU oracle (test infrastructure only) into oracle/liboracle.so.

    python oracle/build.py

gcc -O3 -march=x86-64-v3 -std=c99 -fopenmp -ffp-contract=off -fwrapv: the
reference's emitted-C flags (app/CBackend.hs:89, `cc -O3 -march=native
-std=c99 -fopenmp`) with contraction off and signed wraparound stated
explicitly; x86-64-v3 instead of native so the library built here runs on
the GPU host's CPU too.

The reference's C sources for this path cannot be built on their own: its
kernels are emitted by the Haskell compiler (no GHC in this image) and its
runtime header needs the un-vendored SIMDE submodule, so there is no
oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
FLAGS = ["-O3", "-march=x86-64-v3", "-std=c99", "-fopenmp", "-ffp-contract=off", "-fwrapv",
         "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas"]


def build(verbose=False):
    src = os.path.join(HERE, "hyd_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + FLAGS + ["-o", tmp, src, "-lm"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + p.stdout)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
