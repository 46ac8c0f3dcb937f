import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests' device. Fails (not skips) when the library or GPU is
    missing: a GPU test must never pass on a fallback."""
    import torch
    from strata_amd import _lib
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    L = _lib.lib()
    assert L.strata_init(0) == 0, _lib.last_error()
    return torch.device("cuda:0")
