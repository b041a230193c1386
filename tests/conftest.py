import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI libraries)")


@pytest.fixture(scope="session")
def paths():
    """The hot-path C-ABI library, initialised on device 0.  On a GPU box a
    missing or broken library is an error, never a skip."""
    import kokkos_amd as K

    P = K.init(0)
    return P


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O
