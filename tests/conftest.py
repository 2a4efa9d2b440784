import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(__file__))

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as g:
        return {k: g[k] for k in g.files}


# Cases stored in the golden file: (no, nu, seed)
GOLDEN_CASES = [(3, 5, 7), (4, 6, 11), (5, 7, 3), (2, 9, 5), (6, 4, 21),
                (4, 36, 1001), (4, 87, 1002), (8, 120, 1003)]
SMALL_CASES = GOLDEN_CASES[:5]
