import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libtetra_hip.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialise device 0 through the C ABI; every GPU test depends on this."""
    import tetragpu_amd as t

    t.init(0)
    return t
