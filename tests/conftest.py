import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through libpfem.so)")


def gpu_available():
    # the HIP kernel driver node; avoids initialising a runtime just to ask
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def need_gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
