import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def pytest_collection_modifyitems(config, items):
    # long-running GPU tests last, so a failure in them hides nothing else;
    # the allocation-free entry point's test after everything (it shares one
    # device array per GPU with no other entry point, and must pass on a GPU
    # the other tests have used)
    def order(it):
        if "test_zz_" in it.nodeid:
            return 2
        return 1 if it.get_closest_marker("slow") else 0
    items.sort(key=order)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
