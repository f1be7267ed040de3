import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def pytest_collection_modifyitems(config, items):
    # GPU tests report which case runs before touching the device, so that a
    # failing run names its test even if the process dies.
    pass


@pytest.fixture(scope="session")
def gpu_lib():
    """The loaded HIP library (tests marked gpu only)."""
    from fieldspice_amd import _lib
    lib = _lib.load()
    return lib
