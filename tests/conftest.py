import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    return O.lib()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (the reference tree is not on this machine)")
    return O.ref()
