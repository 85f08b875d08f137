import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pytest.skip("oracle not built (make -C oracle)")
    pyoracle.lib()  # a built oracle that does not load is a failure, not a skip
    return pyoracle
