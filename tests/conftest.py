import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libixchel2d.so)")
    config.addinivalue_line("markers", "slow: long CPU run")


def pytest_collection_modifyitems(config, items):
    # print the case before every GPU test so a crash names it
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(pytest.mark.usefixtures("_announce"))


@pytest.fixture
def _announce(request):
    print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
