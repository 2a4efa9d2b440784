"""Write tests/golden/golden.npz from the reference's own compiled routines.

Run in the container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden.py

What it stores (numbers only; the inputs are regenerated from their seeds by
libcc_amd.synth, and a checksum of every input array is stored so a change to
the generator is caught):

* energies (ets, etd) and v1(nu,no) of the full driver for each case in CASES;
* the W cube the reference leaves in V3, and the v1 / etd partials, for single
  tuples of each kind on the (4,6,11) case;
* the 13 cube-operation tables (trant3_1/3/4/5, TRANT3_ACC 1..5, tranmd_23,
  SYMT311_ACC 23/12/13) on the 2x2x2 and 3x3x3 cubes filled with 1..n^3, the
  known-answer table that the reference's tools/trant3 program prints;
* TRANMD_SMP's output for every code on 4-index arrays filled with 1..n.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from libcc_amd import synth  # noqa: E402
from oracle import ref  # noqa: E402

CASES = [(3, 5, 7), (4, 6, 11), (5, 7, 3), (2, 9, 5), (6, 4, 21),
         (4, 36, 1001), (4, 87, 1002), (8, 120, 1003)]
# (kind, i, j, k), 0-based: kind 0 the distinct tuple (i,j,k); kind 1 the tuple (i,j,j) and
# kind 2 the tuple (i,i,j), both from (i, j) with i > j (k unused)
TUPLES = [(0, 2, 1, 0), (0, 3, 2, 1), (0, 3, 1, 0), (1, 2, 1, 1), (1, 3, 0, 0),
          (2, 2, 1, 0), (2, 3, 0, 0)]
# TRANMD_SMP codes on the 3^4 array and on shapes where only the permuted extents agree
CODES = (12, 13, 14, 23, 24, 34, 231, 312, 341, 413, 1234, 99)
TRANMD = [((3, 3, 3, 3), c) for c in CODES] + [
    ((3, 3, 2, 4), 12), ((3, 2, 3, 4), 13), ((3, 2, 4, 3), 14), ((2, 3, 3, 4), 23), ((2, 3, 4, 3), 24),
    ((2, 4, 3, 3), 34), ((3, 3, 3, 2), 231), ((3, 3, 3, 2), 312), ((3, 2, 3, 3), 341), ((3, 2, 3, 3), 413),
    ((3, 2, 2, 3), 1234), ((2, 3, 4, 5), 99)]


def checksum(inp: dict) -> np.ndarray:
    return np.array([float(np.sum(inp[k])) for k in ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")]
                    + [float(np.sum(np.abs(inp[k]))) for k in ("t2", "vm", "voe", "vvvo")])


def main() -> None:
    out = {}
    for no, nu, seed in CASES:
        inp = synth.make_inputs(no, nu, seed)
        r = ref.triples(inp)
        tag = f"{no}_{nu}_{seed}"
        out[f"e_{tag}"] = np.array([r["ets"], r["etd"]])
        out[f"v1_{tag}"] = r["v1"]
        out[f"sum_{tag}"] = checksum(inp)
        print(tag, r["ets"], r["etd"], r["e"])
    inp = synth.make_inputs(4, 6, 11)
    for kind, i, j, k in TUPLES:
        t = ref.tuple_(inp, kind, i, j, k)
        tag = f"{kind}_{i}_{j}_{k}"
        out[f"w_{tag}"] = t["v3"]
        out[f"tv1_{tag}"] = t["v1"]
        out[f"tetd_{tag}"] = np.array([t["etd"]])
    for n in (2, 3):
        out[f"cube_{n}"] = ref.cube_tables(n)
    for shape, code in TRANMD:
        out["tranmd_%d_%d_%d_%d_%d" % (shape + (code,))] = ref.tranmd_table(shape, code)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
