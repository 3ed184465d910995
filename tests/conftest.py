import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    d = os.path.join(ROOT, "tests", "golden")

    def load(name):
        return np.load(os.path.join(d, name + ".npz"), allow_pickle=False)

    return load


@pytest.fixture(scope="session")
def hip():
    """One HIP backend context for the GPU tests (fails loudly without the library or a gfx950 GPU)."""
    from image360_amd import HipComputeBackend

    be = HipComputeBackend(0)
    rc = be.Initialize()
    assert rc == 0, f"HipComputeBackend.Initialize failed: {rc} {be.GetLastError()}"
    yield be
    be.Shutdown()
