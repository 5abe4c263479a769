import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "youtube-simulations_amd", os.path.join("tests", "golden")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def say(*a):
    """progress line naming the case, flushed before the GPU work starts"""
    print(*a, flush=True)
    sys.stdout.flush()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and libwavehip.so")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle
