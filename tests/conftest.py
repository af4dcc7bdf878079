import os
import sys

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, TESTS)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, os.path.join(os.path.dirname(TESTS), "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before it touches the GPU, so a failing run says which case it was.
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu case] {request.node.nodeid}", flush=True)
    yield
