import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import orc
    if not os.path.exists(orc.LIB):
        orc.build()
    return orc.lib()


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when the HIP library or the GPU is missing (no CPU fallback)."""
    import torch  # noqa: F401  (initialises the ROCm runtime the library also uses)
    from k3dom_amd import _lib
    lib = _lib.load()
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    return lib
