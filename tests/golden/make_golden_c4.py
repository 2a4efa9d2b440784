"""Write tests/golden/golden_c4.npz: per-tuple results of the reference's own compiled routines
at the full no=24, nu=216 shape (seed 1004), where the whole driver is too slow to use as a
test oracle.

Run in the container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden_c4.py

For each tuple in TUPLES the harness runs one DDCC_T_IJK_ACC (distinct), DDCC_T_IJJ_ACC or
DDCC_T_IIJ_ACC call (src/triples-ijk/ddcc_t_ijk_big.f:1-21, ddcc_t_ijj_iij_big.f:18-94) on
the seeded generator's arrays (libcc_amd.synth.make_inputs).  Stored, numbers only:

* ``etd_<tag>``: the E[T] the call adds to the module accumulator;
* ``v1_<tag>``: the three v1 columns (nu x 3, columns i, j, k of v1(nu,no)) it adds;
* ``widx_<tag>`` / ``w_<tag>``: for distinct tuples, 256 seeded (a,b,c) positions (0-based)
  and the W_ijk values the reference leaves there in V3 (W(a,a,a) = 0 after zerot3);
* ``sum``: the input checksum of make_golden.checksum, so a change of the generator shows.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
sys.path.insert(0, HERE)

from libcc_amd import synth  # noqa: E402
from oracle import ref  # noqa: E402

from make_golden import checksum  # noqa: E402

NO, NU, SEED = 24, 216, 1004
# (kind, i, j, k), 0-based, as in make_golden.TUPLES
TUPLES = [(0, 23, 22, 21), (0, 10, 5, 0), (0, 23, 11, 3), (0, 2, 1, 0), (1, 12, 3, 3), (2, 20, 7, 7),
          (1, 23, 0, 0), (2, 1, 0, 0)]
NSAMPLE = 256


def tag(t) -> str:
    return "%d_%d_%d_%d" % t


def occupied(t) -> tuple:
    kind, i, j, k = t
    return (i, j, k) if kind == 0 else (i, j, j) if kind == 1 else (i, i, j)


def main() -> None:
    inp = synth.make_inputs(NO, NU, SEED)
    packed = synth.pack_vvvo(inp["vvvo"])
    out = {"sum": checksum(inp)}
    rng = np.random.default_rng(424242)
    for t in TUPLES:
        kind, i, j, k = t
        o = ref.run(inp, {0: 2, 1: 3, 2: 4}[kind], i + 1, j + 1, k + 1, packed=packed)
        nu3 = NU ** 3
        etd = o[0]
        v1 = o[1:1 + NU * NO].reshape(NU, NO, order="F")
        occ = occupied(t)
        cols = sorted(set(occ))
        # the call touches only the columns of its occupied indices
        assert np.all(v1[:, [c for c in range(NO) if c not in cols]] == 0.0)
        out["etd_" + tag(t)] = np.array([etd])
        out["v1_" + tag(t)] = np.asfortranarray(v1[:, list(occ)])
        if kind == 0:
            v3 = o[1 + NU * NO:1 + NU * NO + nu3].reshape(NU, NU, NU, order="F")
            idx = rng.integers(0, NU, size=(NSAMPLE, 3))
            idx[:8] = [[a, a, a] for a in rng.integers(0, NU, 8)]          # zerot3 positions
            idx[8:16, 1] = idx[8:16, 0]                                       # W(a,a,c)
            out["widx_" + tag(t)] = idx.astype(np.int32)
            out["w_" + tag(t)] = v3[idx[:, 0], idx[:, 1], idx[:, 2]]
        print(tag(t), occ, etd, float(np.abs(v1).max()), flush=True)
    np.savez_compressed(os.path.join(HERE, "golden_c4.npz"), **out)


if __name__ == "__main__":
    main()
