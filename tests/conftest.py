import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs libpypohip.so kernels)")


@pytest.fixture(scope="session")
def hip():
    """libpypohip.so with a visible device; fails (not skips) when the
    library is missing, so a GPU run can never pass without native code."""
    from pypo_amd import bind as B
    lib = B.loadHIPlib()
    if lib.pypo_hip_device_count() < 1:
        pytest.fail("no HIP device visible to libpypohip.so")
    return lib


@pytest.fixture(scope="session")
def port():
    from oracle import bind as OB
    return OB.load("port")
