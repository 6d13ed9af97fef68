import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def mp_golden():
    return np.load(os.path.join(GOLDEN, "mp_golden.npz"))


@pytest.fixture(scope="session")
def d_golden():
    return np.load(os.path.join(GOLDEN, "d_golden.npz"))


@pytest.fixture(scope="session")
def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
