import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libiclBLAS.so")


@pytest.fixture(scope="session")
def handle():
    """Host-pointer-mode handle (the reference's calling convention)."""
    from clgpu_amd import iclblas as ib
    h = ib.Handle()
    yield h
    h.close()


@pytest.fixture(scope="session")
def dev_handle():
    """Device-pointer-mode handle bound to torch's current stream."""
    import torch
    from clgpu_amd import iclblas as ib
    torch.cuda.init()
    h = ib.Handle(memory_mode=ib.MEMORY_DEVICE, stream=torch.cuda.current_stream().cuda_stream)
    yield h
    h.close()
