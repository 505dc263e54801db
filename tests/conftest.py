import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbgpq.so on cuda:0)")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def gpu():
    """The loaded product library; GPU tests fail loudly if it is missing."""
    import gpuheap_amd
    lib = gpuheap_amd.lib()
    name = gpuheap_amd.device_name()
    print(f"[gpu] {name}", flush=True)
    return lib
