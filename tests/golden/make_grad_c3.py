"""Write tests/golden/grad_c3.npz: the FP64 gradient of E(T) at no=8, nu=120 (seed 1003) from
the dense closed form, differentiated one ordered occupied triple at a time by torch autograd
(oracle.dense.grad_tuplewise; the dense W of this shape would take 7 GB).

    python tests/golden/make_grad_c3.py

Stored, numbers only: ``e`` (E(T)); ``dd`` (8 x 5): the directional derivative of E along
direction n for each array k in KEYS, dd[n, k] = sum grad_k * dir(n, k), with dir(n, k) the
standard normal array of that array's shape drawn from default_rng(DIR_SEED + 16 n + k);
``gnorm`` (5): the Frobenius norm of each gradient; ``idx_<k>`` / ``g_<k>``: 64 seeded flat
(Fortran-order) positions of each gradient and the values there.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from libcc_amd import synth  # noqa: E402
from oracle import dense  # noqa: E402

NO, NU, SEED = 8, 120, 1003
KEYS = ("t1", "t2", "vm", "voe", "vvvo")
DIR_SEED = 7000
NDIR = 8
NSAMPLE = 64


def direction(n: int, k: int, shape) -> np.ndarray:
    return np.asfortranarray(np.random.default_rng(DIR_SEED + 16 * n + k).standard_normal(shape))


def main() -> None:
    inp = synth.make_inputs(NO, NU, SEED)
    t0 = time.time()
    g = dense.grad_tuplewise(inp, progress=lambda n: print(f"{n} triples, {time.time() - t0:.0f} s", flush=True)
                             if n % 50 == 0 else None)
    out = {"e": np.array([g["e"]])}
    dd = np.zeros((NDIR, len(KEYS)))
    for n in range(NDIR):
        for q, k in enumerate(KEYS):
            dd[n, q] = float(np.sum(g[k] * direction(n, q, inp[k].shape)))
    out["dd"] = dd
    out["gnorm"] = np.array([np.linalg.norm(g[k]) for k in KEYS])
    rng = np.random.default_rng(8080)
    for k in KEYS:
        flat = g[k].ravel(order="F")
        idx = rng.integers(0, flat.size, NSAMPLE)
        idx[0] = int(np.argmax(np.abs(flat)))
        out["idx_" + k] = idx
        out["g_" + k] = flat[idx]
    np.savez_compressed(os.path.join(HERE, "grad_c3.npz"), **out)
    print("E", g["e"], "norms", out["gnorm"], flush=True)


if __name__ == "__main__":
    main()
