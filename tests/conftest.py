import gc
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(autouse=True)
def _release_device_memory(request):
    """After every GPU test: molecules and their XC plans reference each other, so what a test
    leaves behind -- including the library's own device workspaces of batched plans -- is freed
    only by the cyclic collector; run it (and return torch's cached blocks) before the next
    test allocates."""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    gc.collect()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.empty_cache()
    except ImportError:
        pass
