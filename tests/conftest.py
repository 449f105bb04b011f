import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def device():
    """Binds the library to device 0 (the HIP path; there is no fallback)."""
    from reffine_amd import runtime
    runtime.init(0)
    return runtime.device_info()
