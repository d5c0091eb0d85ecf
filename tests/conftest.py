import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU (MI355X) and the built engine")
    config.addinivalue_line(
        "markers", "slow: long CPU statistics runs")


@pytest.fixture(scope="session")
def engine():
    """One engine for the GPU tests of a session (fails loudly when the
    library or the device is missing: there is no CPU fallback)."""
    from ion_gym_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()
