"""Runs the reference's own per-tuple Fortran routines (oracle/_ref/ref_triples).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The binary is compiled
from the reference sources by ``make -C oracle ref`` in the container that has
the reference tree; the prebuilt binary is what other machines use.
"""

from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BINARY = os.path.join(_HERE, "_ref", "ref_triples")


def available() -> bool:
    return os.path.exists(BINARY)


def _pack_vvvo(vvvo: np.ndarray) -> np.ndarray:
    nu, _, _, no = vvvo.shape
    p, q = np.tril_indices(nu)
    return vvvo[p, q, :, :].reshape(len(p), nu * no, order="F")


def run(inp: dict, mode: int, i: int = 0, j: int = 0, k: int = 0, lo: int = 0, hi: int = 0,
        threads: int | None = None, packed: np.ndarray | None = None) -> np.ndarray:
    """Write the input file, run the harness, return the raw float64 output.

    Indices are 1-based here, as the Fortran routines take them."""
    if not available():
        raise RuntimeError(f"{BINARY} is not built; run `make -C oracle ref`")
    no, nu = inp["no"], inp["nu"]
    if packed is None:
        packed = _pack_vvvo(inp["vvvo"])
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([mode, no, nu, i, j, k, lo, hi], dtype=np.int64).tofile(f)
            for key in ("eh", "ep", "t1", "t2", "vm", "voe"):
                np.ravel(inp[key], order="F").astype(np.float64).tofile(f)
            np.ravel(packed, order="F").astype(np.float64).tofile(f)
        env = dict(os.environ)
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(threads)
            env["MKL_NUM_THREADS"] = str(threads)
        subprocess.run([BINARY, fin, fout], check=True, env=env)
        return np.fromfile(fout, dtype=np.float64)


def triples(inp: dict, lo: int = 0, hi: int | None = None, degenerate: bool = True,
            threads: int | None = None, packed: np.ndarray | None = None, deg_pairs: int | None = None) -> dict:
    """Driver over distinct tuples [lo,hi) (0-based ids) plus the degenerate tuples (i,j,j), (i,i,j)
    of the first ``deg_pairs`` pairs i>j in the driver's order (all of them by default; none when
    ``degenerate`` is False)."""
    no, nu = inp["no"], inp["nu"]
    if hi is None:
        hi = no * (no - 1) * (no - 2) // 6
    npairs = no * (no - 1) // 2
    k = 0 if not degenerate else (npairs if deg_pairs is None else int(deg_pairs))
    o = run(inp, 1, k=k, lo=lo, hi=hi, threads=threads, packed=packed)
    return {"ets": o[0], "etd": o[1], "e": o[0] + o[1], "v1": o[2:2 + nu * no].reshape(nu, no, order="F"),
            "seconds": o[2 + nu * no]}


def tuple_(inp: dict, kind: int, i: int, j: int, k: int = 0) -> dict:
    """One tuple, 0-based indices: kind 0 (i,j,k), 1 (i,j,j), 2 (i,i,j)."""
    no, nu = inp["no"], inp["nu"]
    mode = {0: 2, 1: 3, 2: 4}[kind]
    o = run(inp, mode, i + 1, j + 1, k + 1)
    nu3 = nu ** 3
    return {"etd": o[0], "v1": o[1:1 + nu * no].reshape(nu, no, order="F"),
            "v3": o[1 + nu * no:1 + nu * no + nu3].reshape(nu, nu, nu, order="F")}


def cube_tables(n: int) -> np.ndarray:
    """The 13 cube operations applied to the cube filled with 1..n^3 (harness mode 5)."""
    inp = {"no": 1, "nu": n, "eh": np.zeros(1), "ep": np.zeros(n), "t1": np.zeros(n),
           "t2": np.zeros(n * n), "vm": np.zeros(n), "voe": np.zeros(n * n),
           "vvvo": np.zeros((n, n, n, 1))}
    o = run(inp, 5)
    return o.reshape(13, n, n, n).transpose(0, 1, 2, 3).reshape(13, n * n * n)


def tranmd_table(shape: tuple, code: int) -> np.ndarray:
    """a(n1,n2,n3,n4) filled with 1..n1*n2*n3*n4 after the reference's TRANMD_SMP(a,...,code)
    (harness mode 7), as the flat column-major array."""
    inp = {"no": 1, "nu": 1, "eh": np.zeros(1), "ep": np.zeros(1), "t1": np.zeros(1),
           "t2": np.zeros(1), "vm": np.zeros(1), "voe": np.zeros(1), "vvvo": np.zeros((1, 1, 1, 1))}
    n1, n2, n3, n4 = shape
    return run(inp, 7, n1, n2, n3, lo=n4, hi=code)
