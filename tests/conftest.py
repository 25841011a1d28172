import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhotspin_amd.so)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        from oracle import build
        build.build()
    yield


REFERENCE = "/root/reference"


def have_reference():
    return os.path.isdir(os.path.join(REFERENCE, "tests"))
