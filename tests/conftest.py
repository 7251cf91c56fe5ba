import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def gpu():
    from espressopp_amd import _lib
    n = _lib.device_count()
    if n < 1:
        pytest.fail("no GPU visible: gpu tests must run on an MI355X box")
    return n
