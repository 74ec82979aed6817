"""Shared test setup: the gpu marker, paths, and a session-wide library context for GPU tests."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP product library)")


@pytest.fixture(scope="session")
def nttm():
    """ntt_amd with the product library initialised on cuda:0 (GPU tests only)."""
    import torch
    import ntt_amd
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()
    ntt_amd.init(0, tile_log=int(os.environ.get("NTTM_TEST_TILE_LOG", "0")), pts=int(os.environ.get("NTTM_TEST_PTS", "0")))
    yield ntt_amd
    torch.cuda.synchronize()
    ntt_amd.shutdown()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
