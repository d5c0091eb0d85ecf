import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# SDS flights in the tests read ion_gym's own jump table (a fixture under
# tests/golden), as a deployment inside ion_gym does
GOLDEN = os.path.join(ROOT, "tests", "golden")
os.environ.setdefault("ION_GYM_SDS_TABLE",
                      os.path.join(GOLDEN, "sds_jump_icdf.dat"))


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
