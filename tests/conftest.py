import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def eng():
    """The engine's ctypes front end; GPU tests fail (not skip) when the library is absent."""
    import rtscan_amd

    rtscan_amd.lib()
    if rtscan_amd.lib().rts_device_count() < 1:
        pytest.skip("no GPU device visible")
    return rtscan_amd
