"""The hand-written adjoint (ccx_gradient) against independent references.

* dense:  torch autograd (float64, CPU) through the dense closed form of E(T)
  (oracle/dense.py), on small shapes, including padded virtual blocks, nu = 1 and
  degenerate-only tuple sets;
* model:  the per-tuple numpy adjoint (tests/adjoint_model.py);
* FD:     central finite differences of the GPU FP64 energy, step 1e-5 relative to the
  RMS of each input, along 8 seeded random directions at no=8, nu=120 (the bar is a
  relative error <= 1e-6 of the directional derivative);
* fixture: the FP64 dense-einsum gradient at no=8, nu=120 (tests/golden/grad_c3.npz),
  directional derivatives and sampled entries;
* the torch.autograd.Function wrapper, run-to-run bitwise determinism, and that
  splitting the tuple loop into batches changes nothing beyond rounding."""

import os

import numpy as np
import pytest

from libcc_amd import synth
from libcc_amd.triples import TriplesContext
from oracle import dense

import adjoint_model

pytestmark = pytest.mark.gpu

KEYS = ("t1", "t2", "vm", "voe", "vvvo")


def _grad(inp):
    ctx = TriplesContext(inp["no"], inp["nu"])
    try:
        ctx.load(inp)
        return ctx.gradient()
    finally:
        ctx.close()


def _close(got, want, rtol):
    # nu = 1 has only a = b = c points: every gradient is 0 up to the rounding of the six
    # separately accumulated orderings there, hence the floor on the scale
    scale = max(np.abs(want).max(), 1e-20)
    err = np.abs(got - want).max() / scale
    assert err <= rtol, err


@pytest.mark.parametrize("no,nu,seed", [(3, 5, 7), (4, 6, 11), (2, 17, 3), (5, 1, 2), (3, 33, 8), (4, 36, 1001),
                                        (3, 113, 9)],
                         ids=lambda *a: None)
def test_gradient_matches_dense_autograd(no, nu, seed):
    """Padded blocks, nu=1, degenerate-only sets; K = nu + no padded to 4 covers back-contraction
    K chunks of 1, 2 and 3 tiles (kp 8 .. 40) and, at nu=113 (kp=116), a chunk of 8 tiles."""
    inp = synth.make_inputs(no, nu, seed)
    want = dense.grad_dense(inp)
    got = _grad(inp)
    assert abs(got["e"] - want["e"]) <= 1e-12 * max(1.0, abs(want["e"]))
    for k in KEYS:
        _close(got[k], want[k], 1e-11)


def test_gradient_matches_per_tuple_model():
    inp = synth.make_inputs(5, 21, 44)
    want = adjoint_model.model_grad(inp)
    got = _grad(inp)
    for k in ("t2", "vm", "voe", "vvvo"):
        _close(got[k], want[k], 1e-11)


def test_gradient_energy_matches_forward():
    inp = synth.make_inputs(6, 40, 77)
    ctx = TriplesContext(6, 40)
    ctx.load(inp)
    e = ctx.energy()
    g = ctx.gradient()
    ctx.close()
    for k in ("ets", "etd"):
        assert abs(g[k] - e[k]) <= 1e-12 * max(1.0, abs(e[k])), (k, g[k], e[k])
    np.testing.assert_allclose(g["t1"], 2.0 * e["v1"].T, rtol=1e-12, atol=1e-300)


def test_finite_differences_at_adjoint_size():
    """no=8, nu=120: <grad, d> against (E(x+hd) - E(x-hd)) / 2h along 8 random directions."""
    no, nu, seed = synth.CONFIGS["c3"]
    inp = synth.make_inputs(no, nu, seed)
    ctx = TriplesContext(no, nu)
    ctx.load(inp)
    g = ctx.gradient()
    h = 1e-5
    rng = np.random.default_rng(2024)
    worst = 0.0
    for _ in range(8):
        d = {k: rng.standard_normal(inp[k].shape) * np.sqrt(np.mean(inp[k] ** 2)) for k in KEYS}
        lin = sum(float(np.sum(g[k] * d[k])) for k in KEYS)
        es = []
        for sgn in (1.0, -1.0):
            x = dict(inp)
            for k in KEYS:
                x[k] = np.asfortranarray(inp[k] + sgn * h * d[k])
            ctx.load(x)
            es.append(ctx.energy()["e"])
        fd = (es[0] - es[1]) / (2.0 * h)
        rel = abs(fd - lin) / abs(lin)
        worst = max(worst, rel)
        print(f"direction: <g,d> = {lin:.12e}  FD = {fd:.12e}  rel = {rel:.2e}", flush=True)
    ctx.close()
    assert worst <= 1e-6, worst


def test_gradient_matches_dense_fixture_at_adjoint_size():
    """no=8, nu=120 against the FP64 dense-einsum gradient (tests/golden/grad_c3.npz, written by
    tests/golden/make_grad_c3.py through oracle.dense.grad_tuplewise): the directional
    derivative along 8 seeded directions per array, to 1e-10 of |grad| |direction|, and 64
    sampled entries per array, to 1e-11 of the array's max |grad|."""
    from golden.make_grad_c3 import DIR_SEED, NDIR, KEYS as GKEYS, direction  # noqa: F401

    with np.load(os.path.join(os.path.dirname(__file__), "golden", "grad_c3.npz"), allow_pickle=False) as f:
        fix = {k: f[k] for k in f.files}
    no, nu, seed = synth.CONFIGS["c3"]
    inp = synth.make_inputs(no, nu, seed)
    g = _grad(inp)
    assert abs(g["e"] - fix["e"][0]) <= 1e-12 * abs(fix["e"][0])
    worst_dd = worst_s = 0.0
    for q, k in enumerate(GKEYS):
        gn = np.linalg.norm(g[k])
        assert abs(gn - fix["gnorm"][q]) <= 1e-11 * fix["gnorm"][q], k
        for n in range(NDIR):
            d = direction(n, q, inp[k].shape)
            got = float(np.sum(g[k] * d))
            rel = abs(got - fix["dd"][n, q]) / (gn * np.linalg.norm(d))
            worst_dd = max(worst_dd, rel)
            assert rel <= 1e-10, (k, n, got, fix["dd"][n, q])
        flat = g[k].ravel(order="F")[fix["idx_" + k]]
        rel = np.abs(flat - fix["g_" + k]).max() / np.abs(fix["g_" + k]).max()
        worst_s = max(worst_s, rel)
        assert rel <= 1e-11, (k, rel)
    print(f"dense-einsum gradient at no=8 nu=120: directional {worst_dd:.2e}, sampled entries {worst_s:.2e}")


def test_autograd_function_on_device_tensors():
    import torch
    from libcc_amd.autograd import TriplesEnergy
    inp = synth.make_inputs(4, 20, 9)
    dev = {k: torch.tensor(np.ascontiguousarray(inp[k]), dtype=torch.float64, device="cuda",
                           requires_grad=k in KEYS) for k in ("eh", "ep") + KEYS}
    e = TriplesEnergy.apply(dev["eh"], dev["ep"], *(dev[k] for k in KEYS))
    grads = torch.autograd.grad(3.0 * e, [dev[k] for k in KEYS])
    want = _grad(inp)
    assert abs(float(e.detach()) - want["e"]) <= 1e-12 * abs(want["e"])
    for k, gk in zip(KEYS, grads):
        np.testing.assert_allclose(gk.cpu().numpy(), 3.0 * want[k], rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("form", ["fp32", "bf16x3"])
def test_autograd_function_reduced_form(form):
    """The same Function with W recomputed in a reduced form: both the 7-argument call and
    the form argument work, and the gradients stay within 1e-6 of FP64 (Frobenius)."""
    import torch
    from libcc_amd.autograd import TriplesEnergy
    inp = synth.make_inputs(4, 20, 9)
    dev = {k: torch.tensor(np.ascontiguousarray(inp[k]), dtype=torch.float64, device="cuda",
                           requires_grad=k in KEYS) for k in ("eh", "ep") + KEYS}
    e64 = TriplesEnergy.apply(dev["eh"], dev["ep"], *(dev[k] for k in KEYS))
    g64 = torch.autograd.grad(e64, [dev[k] for k in KEYS])
    e = TriplesEnergy.apply(dev["eh"], dev["ep"], *(dev[k] for k in KEYS), form)
    grads = torch.autograd.grad(e, [dev[k] for k in KEYS])
    assert abs(float(e.detach()) - float(e64.detach())) <= 1e-6 * abs(float(e64.detach()))
    for k, gk, gw in zip(KEYS, grads, g64):
        assert float(torch.linalg.norm(gk - gw) / torch.linalg.norm(gw)) <= 1e-6, k


def test_gradient_device_outputs_match_host():
    inp = synth.make_inputs(4, 20, 9)
    ctx = TriplesContext(4, 20)
    ctx.load(inp)
    host = ctx.gradient()
    dev = ctx.gradient(device=True)
    ctx.close()
    for k in KEYS:
        assert np.array_equal(dev[k].cpu().numpy(), host[k]), k


def test_gradient_bitwise_deterministic_and_batched():
    inp = synth.make_inputs(6, 40, 77)
    a, b = _grad(inp), _grad(inp)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    old = os.environ.get("CCX_ADJ_BATCH_MB")
    os.environ["CCX_ADJ_BATCH_MB"] = "1"          # a few tuples per batch
    try:
        c = _grad(inp)
    finally:
        if old is None:
            del os.environ["CCX_ADJ_BATCH_MB"]
        else:
            os.environ["CCX_ADJ_BATCH_MB"] = old
    for k in KEYS:
        _close(c[k], a[k], 1e-13)
