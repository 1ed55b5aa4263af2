import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Tests load what is built; build only if a library is missing."""
    from oracle import build as ob
    if not os.path.exists(ob.LIB):
        ob.build()
    import phafd_amd.build as pb
    if not os.path.exists(pb.LIB):
        pb.build()
    yield
