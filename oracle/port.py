"""ctypes wrapper of oracle/libccport.so (the C restatement of the reference CPU path).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libccport.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle port`")
        L = ctypes.CDLL(path)
        L.port_triples.argtypes = [_i64, _i64] + [_dp] * 7 + [_i64, _i64, ctypes.c_int, _dp, _dp, _dp]
        L.port_triples.restype = ctypes.c_int
        L.port_tuple.argtypes = [_i64, _i64] + [_dp] * 7 + [ctypes.c_int, _i64, _i64, _i64, _dp, _dp, _dp]
        L.port_tuple.restype = ctypes.c_int
        L.port_cube_op.argtypes = [_i64, ctypes.c_int, _dp]
        L.port_cube_op.restype = None
        L.port_task.argtypes = [_i64, _i64, ctypes.POINTER(_i64)]
        L.port_task.restype = None
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(_dp)


def _args(inp):
    return [_p(inp[k]) for k in ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")]


def triples(inp: dict, lo: int = -1, hi: int = -1, degenerate: bool = True) -> dict:
    """ets, etd and v1(nu,no) over distinct tuples [lo,hi) (+ all degenerate ones)."""
    no, nu = inp["no"], inp["nu"]
    v1 = np.zeros((nu, no), order="F")
    ets = ctypes.c_double()
    etd = ctypes.c_double()
    rc = lib().port_triples(no, nu, *_args(inp), lo, hi, int(degenerate), _p(v1),
                            ctypes.byref(ets), ctypes.byref(etd))
    if rc != 0:
        raise MemoryError("port_triples failed")
    return {"ets": ets.value, "etd": etd.value, "e": ets.value + etd.value, "v1": v1}


def tuple_(inp: dict, kind: int, i: int, j: int, k: int = 0) -> dict:
    """One tuple (0-based): kind 0 (i,j,k), 1 (i,j,j), 2 (i,i,j). Returns v3, v1, etd."""
    no, nu = inp["no"], inp["nu"]
    v3 = np.zeros((nu, nu, nu), order="F")
    v1 = np.zeros((nu, no), order="F")
    etd = ctypes.c_double()
    rc = lib().port_tuple(no, nu, *_args(inp), kind, i, j, k, _p(v3), _p(v1), ctypes.byref(etd))
    if rc != 0:
        raise MemoryError("port_tuple failed")
    return {"v3": v3, "v1": v1, "etd": etd.value}


def cube_op(cube: np.ndarray, op: int) -> np.ndarray:
    """Permutation id 1..5 or SYMT311 id 23/12/13 applied to a copy of an n^3 cube."""
    c = np.asfortranarray(cube.astype(np.float64, copy=True))
    lib().port_cube_op(c.shape[0], op, _p(c))
    return c


def task(t: int, no: int) -> tuple:
    out = (_i64 * 3)()
    lib().port_task(t, no, out)
    return tuple(int(x) for x in out)
