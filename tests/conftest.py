import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libmonolis_hip.so")


@pytest.fixture(scope="session")
def hip():
    """The loaded HIP library; a missing build is an error, never a skip."""
    from monolis_amd import _lib
    L = _lib.load()
    assert L.monolis_hip_init() == 0, L.monolis_hip_last_error()
    return L
