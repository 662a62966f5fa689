import os
import sys

import numpy as np
import pytest

# torch ships its own HIP runtime (libamdhip64.so without a SONAME).  Loading
# it before libchx.so makes libchx bind to that same runtime instance, so
# device pointers and streams from torch can be handed to the C ABI.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "chime-x_amd"))

GOLDEN = os.path.join(REPO, "tests", "golden", "reference_vectors.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_cases(golden):
    names = sorted({k.split("__")[0] for k in golden if k.endswith("__params")})
    return names
