import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblgf_flow.so on the GPU)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work starts, so a fault can be attributed
    if request.node.get_closest_marker("gpu"):
        print("\n[case] %s" % request.node.nodeid, flush=True)
    yield
