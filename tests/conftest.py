import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hiplib():
    """The HIP library, built in-tree if it is missing."""
    from hormone_amd import native
    if not os.path.exists(native.LIB_PATH):
        from hormone_amd import build
        build.build()
    return native.lib()
