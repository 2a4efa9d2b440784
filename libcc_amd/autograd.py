"""torch.autograd.Function over the library's hand-written adjoint (checking only).

``TriplesEnergy.apply(eh, ep, t1, t2, vm, voe, vvvo[, form])`` returns E(T) = E[T] + E_ST as
a float64 scalar tensor; ``backward`` returns grad_output times the gradients that
``ccx_gradient_form`` computed in the same call (T2 and VOE independent, VVVO unpacked).
``form`` ("fp64" by default, "fp32" or "bf16x3") is the arithmetic of the W recomputation;
everything after it is FP64.
Orbital energies get no gradient.  Inputs are float64 tensors in the reference
layouts (Fortran order as numpy would give them, i.e. shapes t1 (no,nu), t2
(nu,nu,no,no), vm (no,nu,no,no), voe (nu,nu,no,no), vvvo (nu,nu,nu,no)), on the CPU
or the GPU.  The energy itself is the FP64 GPU path; there is no CPU fallback.
"""

from __future__ import annotations

import numpy as np
import torch

from .triples import TriplesContext

_KEYS = ("t1", "t2", "vm", "voe", "vvvo")


def _fortran(t: torch.Tensor) -> np.ndarray:
    return np.asfortranarray(t.detach().cpu().numpy())


class TriplesEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eh, ep, t1, t2, vm, voe, vvvo, form="fp64"):
        no, nu = eh.shape[0], ep.shape[0]
        tctx = TriplesContext(no, nu)
        try:
            tctx.load({"eh": _fortran(eh), "ep": _fortran(ep), "t1": _fortran(t1), "t2": _fortran(t2),
                       "vm": _fortran(vm), "voe": _fortran(voe), "vvvo": _fortran(vvvo)})
            r = tctx.gradient(form=form)
        finally:
            tctx.close()
        grads = [torch.from_numpy(np.ascontiguousarray(r[k])).to(device=t1.device) for k in _KEYS]
        ctx.save_for_backward(*grads)
        return torch.tensor(r["e"], dtype=torch.float64, device=t1.device)

    @staticmethod
    def backward(ctx, g):
        grads = ctx.saved_tensors
        return (None, None) + tuple(g * x for x in grads) + (None,)
