import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    if not pyoracle.available("port"):
        pytest.skip("oracle/liboracle.so not built (run `make -C oracle`)")
    return pyoracle
