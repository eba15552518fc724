"""Shared test set-up: the repository root on sys.path, the `gpu` marker, and a flushed banner per GPU test so a
failing GPU run names its case."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libgtpeec.so")


@pytest.fixture(autouse=True)
def _announce(request):
    """Print the test's id before a GPU test runs (flushed), so the last line of a crashed run names it."""
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
