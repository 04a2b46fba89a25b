import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pixel-art-remaster-gpu_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs on an MI355X (HIP device required)")


@pytest.fixture(scope="session")
def gpu():
    """The torch module with cuda:0 usable; fails (never skips) without a GPU."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch


@pytest.fixture(scope="session")
def D():
    import depixel

    return depixel


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle
