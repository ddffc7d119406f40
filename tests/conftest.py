import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def gpu_ready():
    """Fails loudly when the HIP library or the device is missing."""
    from rtdl_amd import mi355x_runtime as m

    info = m.mi355x_context_probe()
    print("device:", info, flush=True)
    return info
