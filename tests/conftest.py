import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden_step0():
    import json

    import numpy as np

    meta = json.load(open(os.path.join(GOLDEN, "ref_step0.json")))
    arrays = dict(np.load(os.path.join(GOLDEN, "ref_step0.npz")))
    return meta, arrays


@pytest.fixture(scope="session")
def golden_traj():
    import numpy as np

    return dict(np.load(os.path.join(GOLDEN, "ref_traj.npz")))
