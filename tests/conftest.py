import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ctx():
    import segmentctscans_amd as S
    c = S.make_context(int(os.environ.get("SCT_DEVICE", "0")))
    yield c


@pytest.fixture(scope="session")
def c1_phantom():
    """Config 1: 64^3, cylinder r = 0.45*64, spheres U[4, 12], seed 1."""
    from segmentctscans_amd.phantom import make_phantom
    return make_phantom(64, 64, 4.0, 12.0, seed=1)
