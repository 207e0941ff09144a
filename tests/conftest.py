import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _has_gpu():
    try:
        from universal_classical_montecarlo_amd import abi
        return abi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a gpu-marked test on a box without a device fails loudly."""
    from universal_classical_montecarlo_amd import abi
    n = abi.device_count()
    assert n > 0, "no HIP device visible"
    return 0
