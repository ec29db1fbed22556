import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "photonprop.jl_amd")
for p in (PKG_DIR, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libphotonprop.so)")
    config.addinivalue_line("markers", "slow: longer CPU-side statistical checks")


@pytest.fixture(scope="session")
def ctx():
    """A device context on GPU 0 with the seawater-like medium."""
    import photonprop as pp

    c = pp.Context(pp.SeawaterConfigMedium(), device=0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def medium():
    import photonprop as pp

    return pp.SeawaterConfigMedium()
