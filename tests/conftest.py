import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) device and liblofar_hip.so")


@pytest.fixture(scope="session")
def cuda():
    """The HIP device; GPU tests fail loudly (no silent skip) when it is missing."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    from lofar_amd import lib
    lib()  # raises LibraryMissing if the HIP library is not built
    return torch.device("cuda:0")
