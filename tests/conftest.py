import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


def _ensure_built():
    from cuda_multiddm_amd import engine
    from oracle import oracle
    if not os.path.exists(engine.LIB_PATH) or not os.path.exists(engine.CLI_PATH):
        from cuda_multiddm_amd import build
        build.build()
    oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when the GPU tests run without a HIP device."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")
