import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is present; the HIP library must be built."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ttc_amd import runtime
    runtime.lib()  # fails loudly if libttc_hip.so is missing
    return torch.device("cuda:0")
