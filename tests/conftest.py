import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "aeromhd-cuda_amd")
ORACLE = os.path.join(REPO, "oracle")
for p in (ORACLE, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X / gfx950) and libaeromhd_hip.so")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.build()
    return oracle


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def torch_gpu():
    """torch with a visible GPU; GPU tests fail (not skip) without one."""
    import torch

    assert torch.cuda.is_available(), "gpu-marked test run without a visible GPU"
    import amhd_lib

    amhd_lib.load()
    return torch
