import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgastar.so on the device)")


@pytest.fixture(scope="session")
def engine():
    """One engine per session on device 0 (GPU tests only)."""
    import cuda_a_star_amd as ga

    e = ga.Engine(0, max_iterations=2_000_000)
    yield e
    e.close()
