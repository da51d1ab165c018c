import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (gfx950) device")


@pytest.fixture(scope="session")
def gat():
    import graphautotune_amd as gat
    gat.core.lib()  # raises if libgat.so is missing
    return gat


@pytest.fixture(scope="session")
def orc():
    from oracle import api
    api.lib()
    return api


@pytest.fixture(scope="session")
def gpu(gat):
    if gat.device_count() < 1:
        pytest.fail("GPU test collected but no HIP device is visible")
    return True
