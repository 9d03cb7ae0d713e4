import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "block-dda_py_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden_h5():
    import numpy as np
    d = np.load(os.path.join(GOLDEN, "pcas_ocbs_simulated_data.npz"))
    return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def golden_nb():
    import numpy as np
    d = np.load(os.path.join(GOLDEN, "test_dda_notebook.npz"))
    return {k: d[k] for k in d.files}
