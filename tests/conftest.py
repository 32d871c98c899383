import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    return pyoracle.load()


@pytest.fixture(scope="session")
def hip():
    """The HIP library; fails loudly when it is missing or no GPU is visible."""
    import torch  # noqa: F401  (initialises the ROCm runtime the same way the runner does)
    from sm_2d_amd import runtime
    return runtime.load_library()
