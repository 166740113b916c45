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


@pytest.hookimpl(tryfirst=True)
def pytest_runtest_setup(item):
    # name every GPU test on the terminal before it runs, with output
    # capturing suspended and the line flushed, so that the tail of a run
    # that took the device down says which test it was
    if item.get_closest_marker("gpu") is None:
        return
    capman = item.config.pluginmanager.getplugin("capturemanager")
    line = f"\n[gpu-test] {item.nodeid}\n"
    if capman is None:
        sys.stderr.write(line)
        sys.stderr.flush()
        return
    with capman.global_and_fixture_disabled():
        sys.stderr.write(line)
        sys.stderr.flush()
