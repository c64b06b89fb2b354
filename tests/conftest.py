import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle (test infrastructure).  Built here if absent."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    O.lib()
    return O


@pytest.fixture(scope="session")
def hip_lib():
    """The product library; the GPU tests fail loudly if it is absent."""
    from mglc_amd import lbm
    return lbm.load_library()
