import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")


@pytest.fixture(scope="session")
def hip_device():
    """Skip-free gate: a gpu-marked test on a machine without a device fails
    loudly instead of silently passing on a fallback."""
    from pyrit_amd import device

    if not device.gpu_available():
        pytest.fail("no HIP device visible: gpu tests must run on the MI355X")
    return True
