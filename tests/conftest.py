import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built library")


def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The engine on cuda:0. A GPU test that runs without a GPU or without the
    built library fails loudly: there is no fallback to silently pass on."""
    import torch
    assert torch.cuda.is_available(), "GPU test started without a visible GPU"
    import parparpar_amd as P
    P.lib()
    torch.cuda.init()
    return P
