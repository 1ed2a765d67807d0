import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libradhip.so")


@pytest.fixture(autouse=True)
def _fresh_registry():
    from radiacuda_amd import rad
    rad.UtiDelAll()
    yield
    rad.UtiDelAll()
