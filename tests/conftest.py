import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfpho_hip.so)")


@pytest.fixture(scope="session")
def hip():
    """Initialise the HIP device through the C ABI (GPU tests only)."""
    from forcepho_amd import _lib
    _lib.init(0)
    return _lib
