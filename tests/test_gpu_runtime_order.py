"""The package can be used before torch in a fresh process: libcc_amd.lib() loads torch (and
its HIP runtime) before the library, so torch still finds the GPU afterwards."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
from libcc_amd import synth                     # no torch import by the caller
from libcc_amd.triples import TriplesContext
ctx = TriplesContext(3, 17)
ctx.load(synth.make_inputs(3, 17, 6))
e = ctx.energy("fp64")["e"]
ctx.close()
import torch
x = torch.ones(4, dtype=torch.float64, device="cuda:0")
print("ok", e, float(x.sum()), flush=True)
"""


def test_library_first_then_torch():
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ok" in r.stdout and r.stdout.split()[-1] == "4.0", r.stdout
