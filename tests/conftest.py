import os
import sys

import pytest

# torch first: it loads its own HIP runtime libraries, and libobvh.so and the
# plugins (built against the system ROCm) must find them already in the
# process when a test loads them through ctypes -- loading a plugin before
# torch made later device allocations inside the plugin fail on the GPU.
import torch  # noqa: E402,F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle
