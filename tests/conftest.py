import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hydrangea-compiler_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhydhip kernels)")


@pytest.fixture(scope="session")
def hyd():
    """libhydhip on the GPU (fails loudly when the library or the GPU is missing)."""
    import hydhip

    hydhip.init()
    info = hydhip.device_info()
    print("device:", info, flush=True)
    return hydhip


@pytest.fixture(scope="session")
def orc():
    import pyoracle

    pyoracle.lib()
    return pyoracle
