import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libopx.so")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail (never skip) when the library or the device is missing."""
    from openpixi_amd import _lib

    _lib.load()
    return True
