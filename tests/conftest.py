import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libek_hip.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count only; no GPU initialisation)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
