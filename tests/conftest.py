import os
import sys

import pytest

# PyTorch ships its own HIP / HSA runtime.  Imported before libsbh_hip.so is
# loaded, it makes the library bind to that same HSA runtime; in the other
# order the process holds two runtimes and torch finds no GPU.  Tests that
# hand torch device buffers to the library rely on this order.
try:
    import torch  # noqa: F401
except Exception:  # torch is optional for the CPU tests
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stochastic-barnes-hut_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to ask the runtime whether a GPU is visible)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def sbh_gpu():
    """The sbh_gpu plugin; with a GPU visible it must be the native HIP build."""
    import sbh_gpu as m
    assert m.device_count() >= 1, "no GPU visible to libsbh_hip.so"
    return m
