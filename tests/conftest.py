import os
import sys

import numpy as np
import pytest

# torch (used by a few GPU tests for device-resident inputs) bundles its own HIP runtime and
# links it by the unversioned name; importing it before libcc_mi355x.so is loaded makes the
# library bind to that same runtime (matching soname) instead of a second copy.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

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


@pytest.fixture(scope="session")
def progress(request):
    """say(msg): a progress line on the terminal even under pytest -q with output captured, so a
    test that runs for minutes is never silent for long."""
    capman = request.config.pluginmanager.getplugin("capturemanager")

    def say(msg: str) -> None:
        if capman is None:
            print(msg, flush=True)
            return
        with capman.global_and_fixture_disabled():
            print(msg, flush=True)

    return say
