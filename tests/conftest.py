import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def have_gpu():
    # the kernel driver node exists only where an AMD GPU is visible
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.skip("no GPU visible")
    return 0
