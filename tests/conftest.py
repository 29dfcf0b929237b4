import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "jump-cannon_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def pytest_runtest_setup(item):
    # Name the case before any GPU work starts, so a run that dies mid-test
    # still says which test it was.
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu-test] {item.nodeid}\n")
        sys.stdout.flush()
