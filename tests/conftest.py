import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "distributed-gemm-thesis_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def dev():
    """The GPU under test; fails (does not skip) when there is none."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    import dgamd
    assert dgamd.lib().dg_device_ok() == 1, "device is not gfx950: " + dgamd.last_error()
    return torch.device("cuda", 0)
