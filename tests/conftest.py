import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

DECKS = os.path.join(ROOT, "tests", "golden", "decks")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running (minutes)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def hiplib():
    """The built HIP library; builds it in-tree when nothing is built yet."""
    import tealeaf_opencl_amd as tl
    from tealeaf_opencl_amd.build import build, lib_path
    if not os.path.exists(lib_path()):
        build()
    tl.load_library()
    return tl


@pytest.fixture(scope="session")
def decks_dir():
    return DECKS
