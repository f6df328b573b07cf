import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def lib():
    import labelpropcuda_amd as L
    return L.load()


@pytest.fixture(scope="session")
def c1_inputs():
    """Config 1 of the benchmark: 5,000 vertices, degree 5, 10 batches, seed 42."""
    import labelpropcuda_amd as L
    return L.generate(5000, 5, 10, 42)


@pytest.fixture(scope="session")
def gpu_available():
    import labelpropcuda_amd as L
    if L.device_count() < 1:
        pytest.fail("GPU test collected but no GPU is visible")
    return True
