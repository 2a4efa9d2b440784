"""Parity at the full no=24, nu=216 shape, where the oracle is too slow to run, through
properties the mathematics fixes:

* W is linear in (VVVO, VM) at fixed T2, and scaling by 2 is exact in every FP64 product and
  sum, so doubling VVVO and VM must give E[T] x 4 and E_ST x 2 and v1 x 2 **bitwise**;
* E(T) is a sum over all occupied tuples and all virtual triples, so relabelling the occupied
  or the virtual orbitals (consistently in every array) leaves it unchanged up to summation
  order, and permutes v1 the same way.  Relabelling the virtuals moves every element to
  other 16-blocks and block orbits, so it exercises the whole block-orbit bookkeeping.

The reduced forms (FP32, split BF16x3) are held to the same two properties at this shape:
their power-of-two scaling is bitwise too (FP32 products of exactly doubled operands, and the
split form's slab scales are powers of two, so its slices are exactly halved or kept), and
they agree with the FP64 form to 1e-9 relative on E(T).  Their adjoints (W recomputed in the
reduced form, the rest FP64) are within 1e-6 of the FP64 gradients in the relative Frobenius
norm of every array.

Inputs are the seeded generator's arrays built on the device (libcc_amd.synth)."""

import numpy as np
import pytest

from libcc_amd import synth
from libcc_amd.triples import TriplesContext

pytestmark = pytest.mark.gpu

NO, NU, SEED = 24, 216, 1004


@pytest.fixture(scope="module")
def full():
    import torch

    inp = synth.make_inputs_torch(NO, NU, SEED, "cuda:0")
    ctx = TriplesContext(NO, NU)
    ctx.load(inp)
    ref = ctx.energy()
    yield torch, inp, ctx, ref
    ctx.close()
    del inp
    torch.cuda.empty_cache()


def _c(t, fshape):
    """C-order view of a Fortran-ordered 1-D buffer: axis 0 is the last Fortran index."""
    return t.view(tuple(reversed(fshape)))


def test_power_of_two_scaling_is_exact(full):
    torch, inp, ctx, ref = full
    sc = dict(inp, vvvo=inp["vvvo"] * 2.0, vm=inp["vm"] * 2.0)
    ctx.load(sc)
    r = ctx.energy()
    ctx.load(inp)
    assert r["etd"] == 4.0 * ref["etd"]
    assert r["ets"] == 2.0 * ref["ets"]
    assert np.array_equal(r["v1"], 2.0 * ref["v1"])


def _relabelled(torch, inp, p_occ, p_vir):
    no, nu = NO, NU
    po = torch.as_tensor(p_occ, device="cuda:0")
    pv = torch.as_tensor(p_vir, device="cuda:0")
    out = {"eh": inp["eh"][po].contiguous(), "ep": inp["ep"][pv].contiguous()}
    t1 = _c(inp["t1"], (no, nu))[pv][:, po]                                   # (nu, no)
    t2 = _c(inp["t2"], (nu, nu, no, no))[po][:, po][:, :, pv][:, :, :, pv]    # (no, no, nu, nu)
    vm = _c(inp["vm"], (no, nu, no, no))[po][:, po][:, :, pv][:, :, :, po]    # (no, no, nu, no)
    voe = _c(inp["voe"], (nu, nu, no, no))[po][:, po][:, :, pv][:, :, :, pv]  # (no, no, nu, nu)
    vvvo = _c(inp["vvvo"], (nu, nu, nu, no))[po][:, pv][:, :, pv][:, :, :, pv]  # (no, nu, nu, nu)
    for k, v in (("t1", t1), ("t2", t2), ("vm", vm), ("voe", voe), ("vvvo", vvvo)):
        out[k] = v.contiguous().reshape(-1)
    return out


@pytest.mark.parametrize("which", ["occupied", "virtual"])
def test_relabelling_invariance(full, which):
    torch, inp, ctx, ref = full
    rng = np.random.default_rng(7)
    p_occ = rng.permutation(NO) if which == "occupied" else np.arange(NO)
    p_vir = rng.permutation(NU) if which == "virtual" else np.arange(NU)
    rel = _relabelled(torch, inp, p_occ, p_vir)
    ctx.load(rel)
    r = ctx.energy()
    ctx.load(inp)
    del rel
    scale = abs(ref["etd"])
    assert abs(r["etd"] - ref["etd"]) <= 1e-11 * scale, (r["etd"], ref["etd"])
    assert abs(r["ets"] - ref["ets"]) <= 1e-11 * scale, (r["ets"], ref["ets"])
    # v1(a, i) of the relabelled problem = v1(p_vir[a], p_occ[i]) of the original
    want = ref["v1"][np.ix_(p_vir, p_occ)]
    np.testing.assert_allclose(r["v1"], want, rtol=0, atol=1e-10 * np.abs(want).max())


# (relative bar on E(T), bar on v1 as a fraction of max|v1|, whether doubling VVVO and VM doubles
# every slice product exactly): FP32 products of doubled operands are exact doublings and the
# per-element bf16 slices scale with the element.  Not so for the two forms with shared scales:
# the fixed-point BF16x4 digits sit on one grid per slab whose scale the (undoubled) T2 part sets,
# and the fp8 row group of 32 K that straddles K = nu holds VVVO (doubled) next to T2 (not), so its
# exponent need not move with the VVVO part.
FULL_BARS = {"fp32": (1e-9, 1e-5, True), "bf16x3": (1e-9, 1e-5, True), "bf16x4": (1e-10, 1e-6, False),
             "fp8x4": (2e-8, 5e-4, False)}


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "bf16x4", "fp8x4"])
def test_reduced_form_at_full_size(full, form):
    torch, inp, ctx, ref = full
    print(f"full size form={form}", flush=True)
    r = ctx.energy(form)
    e_bar, v_bar, doubles = FULL_BARS[form]
    print(f"  |dE| {abs(r['e'] - ref['e']):.3e} Eh ({abs(r['e'] - ref['e']) / abs(ref['e']):.2e} relative)")
    assert abs(r["e"] - ref["e"]) <= e_bar * abs(ref["e"]), (r["e"], ref["e"])
    np.testing.assert_allclose(r["v1"], ref["v1"], rtol=0, atol=v_bar * np.abs(ref["v1"]).max())
    if not doubles:
        return
    sc = dict(inp, vvvo=inp["vvvo"] * 2.0, vm=inp["vm"] * 2.0)
    ctx.load(sc)
    r2 = ctx.energy(form)
    ctx.load(inp)
    del sc
    assert r2["etd"] == 4.0 * r["etd"]
    assert r2["ets"] == 2.0 * r["ets"]
    assert np.array_equal(r2["v1"], 2.0 * r["v1"])


def test_reduced_adjoints_at_full_size(full):
    torch, inp, ctx, ref = full
    print("full size adjoint fp64", flush=True)
    g64 = ctx.gradient(device=True)
    assert abs(g64["e"] - ref["e"]) <= 1e-12 * abs(ref["e"])
    for form in ("fp32", "bf16x3"):
        print(f"full size adjoint {form}", flush=True)
        g = ctx.gradient(device=True, form=form)
        for k in ("t1", "t2", "vm", "voe", "vvvo"):
            err = float(torch.linalg.norm(g[k] - g64[k]) / torch.linalg.norm(g64[k]))
            assert err <= 1e-6, (form, k, err)
        del g
    del g64
    torch.cuda.empty_cache()
