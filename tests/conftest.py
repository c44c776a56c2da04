import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through libdfrg.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialise libdfrg.so on device 0 (the native library must be built and a GPU visible)."""
    import diffrg_amd as D
    print(f"[gpu] init libdfrg.so from {D.LIB_PATH}", flush=True)
    D.init(0)
    print(f"[gpu] device {D.device_props()}", flush=True)
    yield D


@pytest.fixture(scope="session")
def qp(gpu):
    q = gpu.QuadratureProvider()
    yield q
    q.close()
