import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# strict failure policy for every test process: a GPU-sized call whose GPU path
# fails ends the process instead of passing on the CPU BLAS
os.environ.setdefault("BLASX_CPU_FALLBACK", "0")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")


@pytest.fixture(scope="session")
def bx():
    import blasx_amd

    blasx_amd.lib()
    return blasx_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(bx):
    """The GPU path must be live: a GPU test never passes on a CPU fallback."""
    assert bx.gpu_available(), "libblasx reports no usable GPU"
    return bx
