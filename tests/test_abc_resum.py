"""The virtual-triple re-summation of E[T] that ccx_energy_abc evaluates, checked on the
CPU against the ijk-driven E[T] of the dense oracle (src/triples-ijk/t1wt3.f90:53-102 and
the driver weights, src/triples-ijk/triples.F90:237-300).

For a virtual triple (a,b,c) the occupied cube is V(i,j,k) = W_ijk(a,b,c), and
    X_abc = sum_ijk V(ijk) [8 V(ijk) - 4 (V(ikj) + V(kji) + V(jik)) + 2 (V(kij) + V(jki))] / D,
    E[T] = sum_{a>b>c} X_abc + 1/2 sum_{a>b} (X_abb + X_aab).
The permuted reads follow from the pair-permutation symmetry of W.  The planted-mistake
cases show that a transposition read as a cyclic permutation, or the degenerate triples
counted with weight 1, changes the sum.  Also pins that V(i,i,i) drops out (the abc
prototype zeroes it, src/triples-abc/abc_tuple.f90:77)."""

import numpy as np
import pytest

from libcc_amd import synth
from oracle import dense


def _x_abc(w, eh, ep, wrong_read=False, zero_diag=True):
    """X_abc for every ordered virtual triple from W[i,j,k,a,b,c]."""
    v = np.transpose(w, (3, 4, 5, 0, 1, 2)).copy()        # v[a,b,c,i,j,k] = W_ijk(a,b,c)
    no = w.shape[0]
    if zero_diag:
        i = np.arange(no)
        v[:, :, :, i, i, i] = 0.0
    def p(s):
        return np.einsum("abc" + s + "->abcijk", v)
    ikj = p("kij") if wrong_read else p("ikj")
    br = 8.0 * v - 4.0 * (ikj + p("kji") + p("jik")) + 2.0 * (p("kij") + p("jki"))
    o = eh[:, None, None] + eh[None, :, None] + eh[None, None, :]
    u = ep[:, None, None] + ep[None, :, None] + ep[None, None, :]
    d = o[None, None, None] - u[:, :, :, None, None, None]
    return (v * br / d).sum(axis=(3, 4, 5))


def _etd_abc(x, wdeg=0.5):
    nu = x.shape[0]
    e = 0.0
    for a in range(nu):
        for b in range(a):
            for c in range(b):
                e += x[a, b, c]
    for a in range(nu):
        for b in range(a):
            e += wdeg * (x[a, b, b] + x[a, a, b])
    return e


@pytest.mark.parametrize("no,nu,seed", [(3, 5, 11), (4, 6, 12), (2, 7, 13)])
def test_abc_resummation_equals_ijk(no, nu, seed):
    inp = synth.make_inputs(no, nu, seed)
    r = dense.energies(inp)
    x = _x_abc(r["w"], inp["eh"], inp["ep"])
    got = _etd_abc(x)
    assert abs(got - r["etd"]) <= 1e-12 * abs(r["etd"]), (got, r["etd"])
    # the diagonal of V3 does not contribute
    x_nz = _x_abc(r["w"], inp["eh"], inp["ep"], zero_diag=False)
    assert abs(_etd_abc(x_nz) - got) <= 1e-13 * abs(got)
    # X is symmetric under permutations of the virtual triple, and zero for a = b = c
    np.testing.assert_allclose(x, np.transpose(x, (1, 0, 2)), rtol=1e-11, atol=1e-14 * np.abs(x).max())
    np.testing.assert_allclose(x, np.transpose(x, (2, 1, 0)), rtol=1e-11, atol=1e-14 * np.abs(x).max())
    a = np.arange(nu)
    assert np.abs(x[a, a, a]).max() <= 1e-13 * np.abs(x).max()


def test_planted_mistakes_are_seen():
    inp = synth.make_inputs(3, 6, 14)
    r = dense.energies(inp)
    bad_read = _etd_abc(_x_abc(r["w"], inp["eh"], inp["ep"], wrong_read=True))
    assert abs(bad_read - r["etd"]) > 1e-6 * abs(r["etd"])
    bad_weight = _etd_abc(_x_abc(r["w"], inp["eh"], inp["ep"]), wdeg=1.0)
    assert abs(bad_weight - r["etd"]) > 1e-6 * abs(r["etd"])
