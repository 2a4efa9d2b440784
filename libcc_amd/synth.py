"""Seeded synthetic closed-shell inputs for the (T) triples correction.

The arrays come out in the layouts the reference's drivers take
(``src/triples-ijk/ddcc_t_ijk_big.f:6-10``, ``src/triples-ijk/program.F90:195-218``),
each as a Fortran-ordered (column-major) numpy array so its raw buffer is exactly
what a Fortran caller would pass:

* ``eh(no)``, ``ep(nu)``      orbital energies, occupied / virtual, sorted
* ``t1(no,nu)``               T1(i,a)
* ``t2(nu,nu,no,no)``         T2(a,b,i,j) = t_ji^ab, symmetric T2(a,b,i,j) = T2(b,a,j,i)
* ``vm(no,nu,no,no)``         VM(l,c,k,j) = (ck|jl)        ("VOOO")
* ``voe(nu,nu,no,no)``        VOE(b,c,j,k) = (bj|ck)       ("VVOO")
* ``vvvo(nu,nu,nu,no)``       VVVO(a,b,c,i) = (ab|ci), symmetric in a<->b

Integrals are density fitted: (pq|rs) = sum_Q B[Q,p,q] B[Q,r,s], B symmetric in
(p,q), naux = 3*(no+nu), entries N(0, 0.05^2) drawn on p<=q and mirrored.
Orbitals are ordered occupied first (0..no-1), then virtual.
T2 is MP2-like: T2(a,b,i,j) = VOE(a,b,j,i) / (eps_i + eps_j - eps_a - eps_b).

Draw order from ``numpy.random.default_rng(seed)`` (fixed, so every consumer that
regenerates a configuration gets identical arrays): eps_occ, eps_vir, B, T1.
"""

from __future__ import annotations

import numpy as np

# Configurations (no, nu, seed).  The seed is 1000 + the configuration number.
CONFIGS = {
    "c1": (4, 36, 1001),     # reference CPU-runnable size
    "c2": (4, 87, 1002),     # plain FP64 GPU parity size
    "c3": (8, 120, 1003),    # adjoint size
    "c4": (24, 216, 1004),   # precision / timing size
    "c5": (30, 400, 1005),   # largest single-GPU size
}


def draw_factors(no: int, nu: int, seed: int):
    """Return (eh, ep, B, t1) exactly as drawn; B has shape (naux, n, n)."""
    rng = np.random.default_rng(seed)
    eh = np.sort(rng.uniform(-1.5, -0.3, no))
    ep = np.sort(rng.uniform(0.1, 3.0, nu))
    n = no + nu
    naux = 3 * n
    b = rng.normal(0.0, 0.05, (naux, n, n))
    upper = np.triu(b)
    b = upper + np.transpose(np.triu(b, 1), (0, 2, 1))
    t1 = rng.normal(0.0, 1e-3, (no, nu))
    return eh, ep, b, t1


def make_inputs(no: int, nu: int, seed: int) -> dict:
    """Build the six reference-layout arrays (numpy, float64, Fortran order).

    Only meant for sizes whose VVVO fits comfortably in host memory; the
    large configurations are generated on the device from the same factors.
    """
    eh, ep, b, t1 = draw_factors(no, nu, seed)
    bvv = b[:, no:, no:]
    bvo = b[:, no:, :no]
    boo = b[:, :no, :no]
    vvvo = np.einsum("Qab,Qci->abci", bvv, bvo, optimize=True)
    vm = np.einsum("Qck,Qjl->lckj", bvo, boo, optimize=True)
    voe = np.einsum("Qbj,Qck->bcjk", bvo, bvo, optimize=True)
    d2 = (eh[None, None, :, None] + eh[None, None, None, :]
          - ep[:, None, None, None] - ep[None, :, None, None])
    t2 = np.transpose(voe, (0, 1, 3, 2)) / d2
    f = np.asfortranarray
    return {
        "no": no, "nu": nu, "seed": seed,
        "eh": f(eh), "ep": f(ep), "t1": f(t1),
        "t2": f(t2), "vm": f(vm), "voe": f(voe), "vvvo": f(vvvo),
    }


def make_inputs_torch(no: int, nu: int, seed: int, device="cuda:0") -> dict:
    """The same arrays as make_inputs, built on the GPU (for sizes whose VVVO is large).

    Each tensor is 1-D and holds the Fortran-ordered buffer of the reference layout, so
    its data pointer can be handed to the C ABI as is.  The factors are drawn on the
    host with the same generator; only the contractions run on the device (summation
    order differs from numpy's einsum at the level of rounding)."""
    import torch

    eh, ep, b, t1 = draw_factors(no, nu, seed)
    tb = torch.from_numpy(b).to(device)
    del b
    bvv = tb[:, no:, no:]
    bvo = tb[:, no:, :no]
    boo = tb[:, :no, :no]
    naux = tb.shape[0]
    # VVVO(a,b,c,i) column-major == C-order [i][c][b][a]
    left = bvo.permute(0, 2, 1).reshape(naux, no * nu)            # (Q, i*c)
    right = bvv.permute(0, 2, 1).reshape(naux, nu * nu)           # (Q, b*a)
    vvvo = (left.t() @ right).reshape(-1)
    del left, right
    vm = torch.einsum("Qck,Qjl->jkcl", bvo, boo).contiguous().reshape(-1)     # VM(l,c,k,j)
    voe4 = torch.einsum("Qbj,Qck->kjcb", bvo, bvo).contiguous()               # VOE(b,c,j,k)
    teh = torch.from_numpy(eh).to(device)
    tep = torch.from_numpy(ep).to(device)
    # T2(a,b,i,j) = VOE(a,b,j,i)/(e_i+e_j-e_a-e_b); C-order [j][i][b][a]
    d2 = (teh[None, :, None, None] + teh[:, None, None, None]
          - tep[None, None, :, None] - tep[None, None, None, :])
    t2 = (voe4.permute(1, 0, 2, 3) / d2).contiguous().reshape(-1)
    del tb
    return {
        "no": no, "nu": nu, "seed": seed,
        "eh": teh, "ep": tep, "t1": torch.from_numpy(np.ascontiguousarray(t1.T)).to(device).reshape(-1),
        "t2": t2, "vm": vm, "voe": voe4.reshape(-1), "vvvo": vvvo,
    }


def make_config(name: str) -> dict:
    no, nu, seed = CONFIGS[name]
    return make_inputs(no, nu, seed)


def full_eri(no: int, nu: int, seed: int) -> tuple:
    """(eh, ep, t1, eri) with eri[p,q,r,s] = (pq|rs) over all orbitals (small sizes only)."""
    eh, ep, b, t1 = draw_factors(no, nu, seed)
    eri = np.einsum("Qpq,Qrs->pqrs", b, b, optimize=True)
    return eh, ep, t1, eri


def pack_vvvo(vvvo: np.ndarray) -> np.ndarray:
    """VVVO(nu,nu,nu,no) -> the reference's distributed layout (nutr, nu*no).

    Column c + nu*(x-1) holds the lower-packed symmetric (p,q) block, row
    index p*(p-1)/2 + q (1-based, q <= p) as read by CPYTSQ /
    DDCC_T_GETVE_acc (``src/triples-ijk/common_functions.f:43-60``,
    ``src/triples-ijk/ddcc_aux.f:18-45``).
    """
    nu, _, _, no = vvvo.shape
    p, q = np.tril_indices(nu)          # row-major lower triangle: (0,0),(1,0),(1,1),...
    packed = vvvo[p, q, :, :]           # (nutr, nu, no)
    return np.asfortranarray(packed.reshape(len(p), nu * no, order="F"))
