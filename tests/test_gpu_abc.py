"""Virtual-triple view (ccx_w_abc, ccx_energy_abc) through the C ABI.

* ccx_w_abc(a,b,c) against the dense closed form of W (oracle/dense.py, the W the 12 GEMMs
  of ijk_tuple build, src/triples-ijk/ijk_tuple.f90:56-84) and against the ijk path's own
  cube (ccx_tuple_w): v3(i,j,k) = W_ijk(a,b,c), v3(i,i,i) = 0 as abc_tuple_formv3 leaves it
  (src/triples-abc/abc_tuple.f90:77).  Triples cross 16-blocks and include equal indices.
* ccx_energy_abc against the ijk-driven etd of ccx_energy (itself pinned to the reference's
  compiled routines): the same E[T] summed over virtual triples, 1e-11 relative."""

import ctypes

import numpy as np
import pytest

from libcc_amd import lib, synth
from libcc_amd.triples import TriplesContext
from oracle import dense

pytestmark = pytest.mark.gpu


def _triples(nu):
    picks = {(1, 1, 1), (nu, 1, 1), (1, nu, 1), (nu, nu, 1), (min(17, nu), 1, nu), (2, 3, 1)}
    rng = np.random.default_rng(nu)
    for _ in range(4):
        picks.add(tuple(int(x) for x in rng.integers(1, nu + 1, size=3)))
    return sorted(picks)


@pytest.mark.parametrize("no,nu,seed", [(3, 17, 6), (4, 36, 1001), (2, 33, 4)], ids=lambda v: str(v))
def test_w_abc_matches_closed_form_and_tuple_w(no, nu, seed):
    inp = synth.make_inputs(no, nu, seed)
    wd = dense.w_dense(inp["t2"], inp["vm"], inp["vvvo"])        # [i,j,k,a,b,c]
    ctx = TriplesContext(no, nu)
    ctx.load(inp)
    scale = np.abs(wd).max()
    cubes = {}
    for (a, b, c) in _triples(nu):
        print(f"w_abc no={no} nu={nu} triple=({a},{b},{c})", flush=True)
        v3 = ctx.w_abc(a, b, c)
        want = wd[:, :, :, a - 1, b - 1, c - 1].copy()
        i = np.arange(no)
        want[i, i, i] = 0.0
        np.testing.assert_allclose(v3, want, rtol=0, atol=1e-12 * scale)
        # the same numbers as the ijk path's cube of each occupied tuple
        for (ii, jj, kk) in [(no, 1, 1), (1, no, no), (no, (no + 1) // 2, 1)]:
            if len({ii, jj, kk}) == 1:
                continue
            key = (ii, jj, kk)
            if key not in cubes:
                cubes[key] = ctx.tuple_w(ii, jj, kk)
            assert abs(v3[ii - 1, jj - 1, kk - 1] - cubes[key][a - 1, b - 1, c - 1]) <= 1e-12 * scale
    ctx.close()


@pytest.mark.parametrize("no,nu,seed", [(4, 36, 1001), (3, 17, 6), (8, 120, 1003), (2, 33, 4), (1, 5, 2), (3, 1, 3)],
                         ids=lambda v: str(v))
def test_energy_abc_matches_ijk(no, nu, seed):
    ctx = TriplesContext(no, nu)
    ctx.load(synth.make_inputs(no, nu, seed))
    print(f"energy_abc no={no} nu={nu}", flush=True)
    want = ctx.energy()["etd"]
    got = ctx.energy_abc()
    assert abs(got - want) <= max(1e-11 * abs(want), 1e-14), (got, want)
    assert ctx.energy_abc() == got          # fixed summation order
    ctx.close()


def test_abc_errors():
    ctx = TriplesContext(2, 5)
    v3 = np.zeros(8)
    et = ctypes.c_double()
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib().ccx_w_abc(ctx._ctx, 1, 1, 1, v3.ctypes.data_as(dp)) == 4        # not loaded
    assert lib().ccx_energy_abc(ctx._ctx, ctypes.byref(et)) == 4
    ctx.load(synth.make_inputs(2, 5, 1))
    assert lib().ccx_w_abc(ctx._ctx, 0, 1, 1, v3.ctypes.data_as(dp)) == 1        # 1-based
    assert lib().ccx_w_abc(ctx._ctx, 1, 6, 1, v3.ctypes.data_as(dp)) == 1
    assert lib().ccx_w_abc(ctx._ctx, 1, 1, 1, None) == 1
    assert lib().ccx_energy_abc(ctx._ctx, None) == 1
    ctx.close()
