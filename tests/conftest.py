import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """Skip cleanly when no GPU is visible; fail loudly when the library is
    missing on a machine that has one."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import fftwpp_amd
    fftwpp_amd.lib()  # raises when libfftwpp_amd.so is not built
    return torch.device("cuda:0")
