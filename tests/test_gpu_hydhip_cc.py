"""The end-to-end program on the GPU: C in the shapes the C backend prints,
with its map-chain and dot loops offloaded, built by hydhip-cc as the HIP
backend builds a program (hydrangea-compiler_amd/_build/e2e/dotmap_program,
made by build()).  Its output against the oracle: the map chain bit for bit
(a checksum of every result's bits), the dot within 1e-12 relative."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "hydrangea-compiler_amd", "_build", "e2e", "dotmap_program")


def expected(orc, n):
    xs = orc.splitmix_f64(n, 17, 0, symmetric=True)
    ys = orc.splitmix_f64(n, 17, n, symmetric=True)
    m = orc.map_chain(xs, ys)
    w = (2 * np.arange(n, dtype=np.uint64) + np.uint64(1))
    with np.errstate(over="ignore"):
        bits = np.bitwise_xor.reduce(m.view(np.uint64) * w) if n else np.uint64(0)
    return orc.dot_dd(m, m), int(bits)


@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_dotmap_program(orc, n):
    assert os.path.exists(EXE), "run build() first"
    print("dotmap_program", n, flush=True)
    r = subprocess.run([EXE, str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fields = dict(kv.split("=") for kv in r.stdout.split())
    dot, bits = float(fields["dot"]), int(fields["bits"], 16)
    ref_dot, ref_bits = expected(orc, n)
    assert bits == ref_bits
    assert abs(dot - ref_dot) <= 1e-12 * abs(ref_dot)
    assert int(fields["threads"]) >= 1
