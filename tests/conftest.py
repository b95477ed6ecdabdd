import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")


def _have_gpu() -> bool:
    try:
        import torch  # noqa: F401  (only to ask the runtime, never for compute)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


HAVE_GPU = _have_gpu()


@pytest.fixture(scope="session")
def gpu():
    if not HAVE_GPU:
        pytest.skip("no GPU visible")
    from mcscram_amd import load_library
    return load_library()
