import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# tests whose input the CPU oracle regenerates in a background process
NEEDS_LARGE24 = {"test_large24_golden_fp64"}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblgf_flow.so on the GPU)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


def pytest_collection_finish(session):
    # start the minutes-long CPU generation of the 24^4 input now, so that it
    # overlaps the tests collected before the one that needs it
    if any(item.name in NEEDS_LARGE24 for item in session.items):
        from tests import large24_input
        large24_input.start()


def pytest_sessionfinish(session, exitstatus):
    from tests import large24_input
    large24_input.cleanup()


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work starts, so a fault can be attributed
    if request.node.get_closest_marker("gpu"):
        print("\n[case] %s" % request.node.nodeid, flush=True)
    yield
