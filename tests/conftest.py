"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on the CPU (oracle, host logic, ABI)."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libplume_hip.so")
    config.addinivalue_line("markers", "slow: long-running")
    config.addinivalue_line("markers", "config: parity at a benchmark configuration's full size (GPU, minutes)")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test ran without a GPU: select it with -m gpu only on a GPU machine")
    from plume_amd.sim.device import init_device

    return init_device("cuda")


@pytest.fixture(scope="session")
def oracle32():
    from oracle import Oracle

    return Oracle("f32")


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle

    return Oracle("f64")
