import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libocean.so on the GPU)")


@pytest.fixture(scope="session")
def gpu():
    from research_amd import ocean
    if ocean.device_count() < 1:
        pytest.fail("no HIP device visible, but a gpu-marked test was selected")
    return ocean
