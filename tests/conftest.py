import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: large problem sizes")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O

    return O.load("orc")


@pytest.fixture(scope="session")
def ref():
    from oracle import oracle as O

    if not O.available("ref"):
        pytest.skip("oracle/_ref not built (needs the reference checkout at build time)")
    return O.load("ref")
