import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import mprgp_amd as M
    n = M.device_count()
    if n <= 0:
        raise RuntimeError("GPU test without a visible HIP device: " + M.last_error())
    c = M.Context(0)
    yield c
    c.close()
