import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libibx.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    """Builds IBTransfer contexts; fails loudly if the HIP library or device is missing."""
    from suguru_amd import IBTransfer

    made = []

    def make(grid):
        t = IBTransfer(grid, 0)
        made.append(t)
        return t

    yield make
    for t in made:
        t.close()
