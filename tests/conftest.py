import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libhpcsort.so")


@pytest.fixture(autouse=True)
def _torch_device_first(request):
    """GPU tests: initialise torch's HIP device before libhpcsort.so.  The
    torch wheel carries its own HIP runtime; when the library's runtime (from
    /opt/rocm) comes up first in a process, torch then finds no device."""
    if request.node.get_closest_marker("gpu") is not None:
        import torch
        if not torch.cuda.is_available():
            pytest.skip("no HIP device")
        torch.zeros(1, device="cuda")
    yield


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
