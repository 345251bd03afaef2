import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line("markers", "slow: long CPU-side runs")


@pytest.fixture(scope="session")
def lib():
    from fltk_sph_ice_amd import load

    return load()


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.load()
