import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsparcgpu.so)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libsparcgpu.so, loaded once; GPU tests fail loudly if it is missing."""
    from sparc_amd import gpu
    return gpu.load()


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle
