import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Skips when no GPU is visible; fails loudly if the HIP library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blasfeo_amd
    blasfeo_amd.lib()  # raises if the library was not built
    return True
