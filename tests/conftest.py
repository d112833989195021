import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhsgen.so")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def ctx():
    # torch ships its own HIP runtime under the same soname as the system
    # one libhsgen.so links; whichever loads first serves the process, and
    # torch cannot start on the other.  Tests that hand torch device buffers
    # to the library therefore start torch's runtime first, as bench.py does.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    from hybridhsdla_amd.api import Context
    c = Context(0)
    yield c
    c.close()


def rel_fro_lower(x, ref):
    """||tril(x - ref)||_F / ||tril(ref)||_F over the lower triangle incl. diagonal."""
    import numpy as np
    d = np.tril(x - ref)
    return float(np.linalg.norm(d) / np.linalg.norm(np.tril(ref)))
