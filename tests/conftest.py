import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblife_hip.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library on a machine with a GPU; fails (never skips) when the
    library is missing or no device is visible, so a GPU run cannot pass on
    a fallback."""
    from life_amd import lib
    lib.load()
    n = lib.device_count()
    assert n > 0, "no HIP device visible"
    return lib
