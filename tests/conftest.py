import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long CPU-side reference runs")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work so a fault can be traced to it
    if request.node.get_closest_marker("gpu"):
        print("\n[gpu-case] %s" % request.node.nodeid, flush=True)
    yield
