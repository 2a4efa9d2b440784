"""The reduced-precision arithmetic forms against the FP64 form of the same library.

FP32 (FP32 operands, FP32 MFMA accumulation) and the split forms (BF16 x 3 per-element slices,
BF16 x 4 fixed-point slices with exact level-0 sums, FP8 e4m3 x 4 slices under per-row 32-K block
exponents; slice products accumulated in FP32, W formed in FP64) all keep the denominators, the
epilogue and every reduction in FP64.  The bars (BARS) follow the bits each form carries: 1e-6
relative on E[T] and E_ST and 1e-5 of max|v1| for FP32 and BF16x3, tighter for BF16x4, looser
for FP8x4 (v1 sums W against T1-sized factors with cancellation, so it keeps fewer correct
digits)."""

import ctypes

import numpy as np
import pytest

from libcc_amd import lib, synth
from libcc_amd.triples import TriplesContext

pytestmark = pytest.mark.gpu

FORMS = ("fp32", "bf16x3", "bf16x4", "fp8x4")
# bars per form, (relative on the energies, of max|v1| on v1), set from what each form carries:
# FP32 and BF16x3 about 24 bits per element, BF16x4 about 27 bits of a typical element with exact
# level-0 sums, FP8x4 16 bits of each 32-element row group
BARS = {"fp32": (1e-6, 1e-5), "bf16x3": (1e-6, 1e-5), "bf16x4": (1e-9, 1e-7), "fp8x4": (3e-5, 5e-4)}


def _close(got, want, rtol, v1_rtol=1e-5):
    # absolute floor 1e-12 Eh: at nu=1 every W element is a diagonal one and the energies are
    # FP64 round-off (1e-23), so there is no relative scale to compare against
    tol = max(rtol * max(abs(want["etd"]), abs(want["ets"])), 1e-12)
    assert abs(got["etd"] - want["etd"]) <= tol, (got["etd"], want["etd"])
    assert abs(got["ets"] - want["ets"]) <= tol, (got["ets"], want["ets"])
    vtol = max(v1_rtol * np.abs(want["v1"]).max(), 1e-12)
    assert np.abs(got["v1"] - want["v1"]).max() <= vtol


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("no,nu,seed", [(4, 36, 1001), (3, 17, 6), (7, 20, 12), (8, 120, 1003), (5, 1, 2),
                                        (2, 33, 4)], ids=lambda v: str(v))
def test_form_matches_fp64(form, no, nu, seed):
    """Configurations with ragged 16-blocks (17, 20, 33), nu=1, degenerate tuples only (no=2)."""
    ctx = TriplesContext(no, nu)
    ctx.load(synth.make_inputs(no, nu, seed))
    want = ctx.energy("fp64")
    got = ctx.energy(form)
    _close(got, want, *BARS[form])
    # run to run bitwise
    again = ctx.energy(form)
    assert again["etd"] == got["etd"] and again["ets"] == got["ets"]
    assert np.array_equal(again["v1"], got["v1"])
    ctx.close()


@pytest.mark.parametrize("form", FORMS)
def test_form_follows_reload(form):
    """The converted operand slabs are rebuilt after every load: a second input set with
    VVVO and VM scaled by 3 (not a power of two, so the split scales change) gives that
    input's energy, not a stale one."""
    no, nu, seed = 6, 40, 77
    inp = synth.make_inputs(no, nu, seed)
    ctx = TriplesContext(no, nu)
    ctx.load(inp)
    first = ctx.energy(form)
    sc = dict(inp, vvvo=np.asfortranarray(inp["vvvo"] * 3.0), vm=np.asfortranarray(inp["vm"] * 3.0))
    ctx.load(sc)
    want = ctx.energy("fp64")
    got = ctx.energy(form)
    _close(got, want, *BARS[form])
    assert abs(got["etd"] - 9.0 * first["etd"]) <= 10 * BARS[form][0] * abs(got["etd"])
    ctx.close()


def test_no_tuples_is_zero():
    ctx = TriplesContext(1, 20)
    ctx.load(synth.make_inputs(1, 20, 3))
    for form in FORMS:
        r = ctx.energy(form)
        assert r["etd"] == 0.0 and r["ets"] == 0.0
    ctx.close()


def test_unknown_form_rejected():
    ctx = TriplesContext(3, 5)
    ctx.load(synth.make_inputs(3, 5, 7))
    ets, etd = ctypes.c_double(), ctypes.c_double()
    st = lib().ccx_energy(ctx._ctx, 7, ctypes.byref(ets), ctypes.byref(etd), None, 0)
    assert st != 0
    with pytest.raises(ValueError):
        ctx.energy("fp16")
    ctx.close()


@pytest.mark.parametrize("form", ("fp32", "bf16x3"))
@pytest.mark.parametrize("no,nu,seed", [(4, 36, 1001), (3, 17, 6), (8, 120, 1003), (2, 33, 4)], ids=lambda v: str(v))
def test_reduced_adjoint_matches_fp64(no, nu, seed, form):
    """ccx_gradient_form(CCX_FORM_FP32 / CCX_FORM_BF16X3): W recomputed in the reduced form,
    dE/dW and every back-contraction in FP64.  Each gradient array within 1e-6 of the FP64
    gradient in the relative Frobenius norm (the bar the FP64 adjoint meets against finite
    differences)."""
    ctx = TriplesContext(no, nu)
    ctx.load(synth.make_inputs(no, nu, seed))
    want = ctx.gradient()
    got = ctx.gradient(form=form)
    assert abs(got["e"] - want["e"]) <= 1e-6 * abs(want["e"])
    for k in ("t1", "t2", "vm", "voe", "vvvo"):
        ref = np.linalg.norm(want[k])
        assert np.linalg.norm(got[k] - want[k]) <= 1e-6 * ref, k
    # the reduced W is the one the energy entry computes
    e = ctx.energy(form)
    assert abs(got["e"] - e["e"]) <= 1e-12 * abs(e["e"])
    again = ctx.gradient(form=form)
    assert again["e"] == got["e"] and np.array_equal(again["vvvo"], got["vvvo"])
    ets, etd = ctypes.c_double(), ctypes.c_double()
    assert lib().ccx_gradient_form(ctx._ctx, 7, ctypes.byref(ets), ctypes.byref(etd), None, None, None, None, None,
                                   0) == 6
    ctx.close()
