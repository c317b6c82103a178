import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def _torch_first():
    """Tests that hand torch tensors to the library need torch's HIP runtime
    initialised before libcfmi.so's (the library does this when torch is
    already imported)."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


_torch_first()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
