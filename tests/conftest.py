import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "oracle"), TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def pkg():
    import importlib
    return importlib.import_module("shmem-cuda-filtration_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent at build time)")
    return oracle.RefCPU()


@pytest.fixture(scope="session")
def hip():
    """The HIP backend through its C ABI, on cuda:0.  GPU tests only."""
    import importlib
    mod = importlib.import_module("shmem-cuda-filtration_amd")
    return mod.device.Backend()
