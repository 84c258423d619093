import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and liboncohip.so")


@pytest.fixture(scope="session")
def ctx():
    from oncolytica_amd.hip import _abi
    c = _abi.Context(0)
    yield c
    c.close()
