import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def engine_factory():
    """Creates engine contexts on device 0; fails (never skips) on a GPU run
    when the HIP library or the device is missing."""
    from jams_amd.engine import Engine

    made = []

    def make():
        e = Engine(0)
        made.append(e)
        return e

    yield make
    for e in made:
        e.close()
