import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def gpu():
    """Fail loudly (never skip) when a GPU test runs without a device."""
    import ltpg_amd

    n = ltpg_amd.device_count()
    assert n > 0, "GPU test selected but no HIP device is visible"
    return n
