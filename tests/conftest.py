import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "toy-path-tracing_amd"
for p in (str(ROOT), str(PKG), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through libtpt on the device)")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    """One libtpt device handle for the whole GPU session."""
    if not _gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
    import tpt

    d = tpt.Device(0)
    yield d
    d.close()
