import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "matrixproductstates.jl_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmpship.so")


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


@pytest.fixture(autouse=True)
def _reset_oracle_tol():
    from oracle import mps_oracle as O
    O.set_qr_rank_tol(0.0)
    yield
    O.set_qr_rank_tol(0.0)


@pytest.fixture(scope="session", autouse=True)
def _torch_hip_first(request):
    """Bring up torch's HIP runtime before the library's when GPU tests run.

    The buffer-level parity helpers (tests/devbuf.py) pass torch device memory to
    the library.  torch ships its own HIP runtime next to the one the library
    links; on the GPU hosts torch then sees no device if the library's runtime
    opened it first, so torch goes first in every test process."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except Exception:
            pass
    yield
