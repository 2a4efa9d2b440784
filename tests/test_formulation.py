"""The GPU kernel's index algebra, replayed in numpy (tests/kernel_model.py), against the
dense closed form: block orbits with the multiplicity weights, the six-ordering
accumulators and their term tables, the Q-form energy and the 36-term v1 epilogue."""

import numpy as np
import pytest

import kernel_model as km
from libcc_amd import synth
from oracle import dense


@pytest.mark.parametrize("pair", [True, False], ids=["paired-modes", "three-ordering-modes"])
@pytest.mark.parametrize("no,nu,seed,tb", [(3, 5, 7, 4), (4, 6, 11, 4), (3, 9, 2, 4), (5, 7, 3, 2), (2, 6, 8, 2)])
def test_kernel_model_matches_dense(no, nu, seed, tb, pair):
    """Ragged blocks (nu not a multiple of tb) take role B; with pair, a degenerate tuple on an
    orbit with equal blocks accumulates two orderings (one on A=B=C) and the kernel's
    completion fills the rest from the cube's transposes and the tuple's pointwise equalities."""
    inp = synth.make_inputs(no, nu, seed)
    m = km.model(inp, tb, pair)
    r = dense.energies(inp)
    assert abs(m["etd"] - r["etd"]) <= 1e-14 * abs(r["etd"])
    assert abs(m["ets"] - r["ets"]) <= 1e-12 * abs(r["ets"]) + 1e-20
    np.testing.assert_allclose(m["v1"], r["v1"], rtol=0, atol=1e-12 * np.abs(r["v1"]).max())


def test_term_tables_cover_each_pairing_once():
    """Each accumulator P receives exactly the six pairings s(P) of the closed form."""
    for p in range(6):
        seen = set()
        for phase in range(3):
            s1, b1, s2, b2 = km.term(phase, p)
            for (s, bq) in ((s1, b1), (s2, b2)):
                x = s % 3
                y, z = [(yy, zz) for yy in range(3) for zz in range(3) if yy != zz and km.pair_index(yy, zz) == bq][0]
                # role of the free virtual index of BX and of the two AX indices
                free = {0: "b", 1: "c", 2: "a"}[phase]
                o = s // 3
                if phase == 0:
                    alpha, beta = ("a", "c") if o == 0 else ("c", "a")
                elif phase == 1:
                    alpha, beta = ("a", "b") if o == 0 else ("b", "a")
                else:
                    alpha, beta = ("b", "c") if o == 0 else ("c", "b")
                pairs = {alpha: x, beta: y, free: z}
                partner = (pairs["a"], pairs["b"], pairs["c"])
                assert partner == km.PERM[p], (p, phase, s, bq)
                seen.add((alpha, beta, free))
        assert len(seen) == 6


def test_q_form_equals_reference_summand():
    """sum over positional permutations of t1wt3's summand = Q(w)."""
    rng = np.random.default_rng(3)
    u = rng.normal(size=(3, 3, 3))          # one orbit: W_ijk at the permutations of (a,b,c)=(0,1,2)
    w = np.array([u[tuple(np.argsort(km.PERM[p]))] for p in range(6)])
    def f(x, y, z):
        d1 = u[x, y, z]
        return d1 * (8 * d1 - 4 * (u[x, z, y] + u[z, y, x] + u[y, x, z]) + 2 * (u[y, z, x] + u[z, x, y]))
    want = sum(f(*km.PERM[s]) for s in range(6))
    ev, od = w[0] + w[3] + w[4], w[1] + w[2] + w[5]
    q = 6 * np.dot(w, w) + 2 * ev * ev + 2 * od * od - 8 * ev * od
    assert abs(q - want) < 1e-12 * abs(want)


def test_adjoint_epilogue_tables_match_per_tuple_adjoint():
    """The item-epilogue tables of the adjoint kernel (Wbar at six placements, the T1.VOE
    products, dVOE partials) against the per-tuple adjoint, with padded 4-blocks."""
    import adjoint_kernel_model
    import adjoint_model
    inp = synth.make_inputs(3, 6, 7)
    want = adjoint_model.model_grad(inp)
    got = adjoint_kernel_model.model_adjoint(inp, tb=4)
    for k in ("vvvo", "t2", "vm", "voe"):
        np.testing.assert_allclose(got[k], want[k], rtol=0, atol=1e-14 * np.abs(want[k]).max())
