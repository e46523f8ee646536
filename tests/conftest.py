import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ffi
    ffi.build()
    return ffi


@pytest.fixture(scope="session")
def hip_backend():
    """The HIP backend on device 0; GPU tests fail loudly when the library is missing."""
    import lumenwright_amd as lw
    b = lw.HipBackend.open(0)
    yield b
    b.close()
