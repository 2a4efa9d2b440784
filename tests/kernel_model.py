"""numpy model of the persistent triples kernel's index algebra (test helper).

It repeats, with plain loops over work items, what libcc_amd/csrc/ccx_kernels.hip
computes: the AX/BX operand slabs, the three K phases with their term tables, the
block-orbit weights, the Q-form energy and the 36-term v1 epilogue.  Comparing it
with the dense oracle checks the formulation on the CPU, independently of the GPU.
"""

from __future__ import annotations

import numpy as np

PERM = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def pair_index(y: int, z: int) -> int:
    return y * 2 + (z - 1 if z > y else z)


def perm_index(a: int, b: int, c: int) -> int:
    return PERM.index((a, b, c))


def term(phase: int, p: int) -> tuple:
    p1, p2, p3 = PERM[p]
    if phase == 0:
        return (p1, pair_index(p3, p2), 3 + p3, pair_index(p1, p2))
    if phase == 1:
        return (p1, pair_index(p2, p3), 3 + p2, pair_index(p1, p3))
    return (p2, pair_index(p3, p1), 3 + p3, pair_index(p2, p1))


def wsrc(p: int, pi) -> int:
    inv = [0, 0, 0]
    for m in range(3):
        inv[pi[m]] = m
    return perm_index(PERM[p][inv[0]], PERM[p][inv[1]], PERM[p][inv[2]])


def wsrc_comp(p: int, s: int, t: int) -> int:
    return wsrc(p, [PERM[s][PERM[t][m]] for m in range(3)])


def tuples(no: int) -> list:
    """kind 1 = distinct, 2 = (i,j,j), 3 = (i,i,j)"""
    out = []
    for i in range(no):
        for j in range(i):
            for k in range(j):
                out.append((i, j, k, 1))
    for i in range(no):
        for j in range(i):
            out.append((i, j, j, 2))
            out.append((i, i, j, 3))
    return out


# orderings each item mode accumulates (the kernel's mode_mask)
MODE_SET = {"full": (0, 1, 2, 3, 4, 5), "aac": (0, 1, 3), "acc": (0, 2, 4), "aaa": (0,),
            "ijj": (0, 2, 3), "iij": (0, 1, 4),
            "ijj_aac": (0, 3), "ijj_acc": (0, 2), "iij_aac": (0, 1), "iij_acc": (0, 4)}


def item_mode(kind: int, mult: int, A: int, B: int, pair: bool = True) -> str:
    """The kernel's mode selection; pair: the forms that combine a degenerate tuple's symmetry
    with equal blocks (pair_modes: FP64, BF16x4, FP8x4)."""
    if kind in (2, 3):
        base = "ijj" if kind == 2 else "iij"
        if mult == 1 or not pair:
            return base
        if mult == 6:
            return "aaa"
        return base + ("_aac" if A == B else "_acc")
    if mult == 1:
        return "full"
    if mult == 6:
        return "aaa"
    return "aac" if A == B else "acc"


def ragged_to_b(nu: int, tb: int, A: int, B: int, C: int, mult: int) -> tuple:
    """The kernel's ragged_block_to_b: a ragged last block of three distinct ones takes role B."""
    if mult == 1 and nu - C * tb < tb:
        return A, C, B
    return A, B, C


def complete(acc: np.ndarray, mode: str) -> np.ndarray:
    """Fill the orderings a mode did not accumulate (the kernel's complete_orderings)."""
    w = acc.copy()
    if mode == "ijj":
        w[1], w[4], w[5] = w[0], w[2], w[3]
    elif mode == "iij":
        w[2], w[3], w[5] = w[0], w[1], w[4]
    elif mode == "ijj_aac":             # w2 = w0(b,a,c); then w1 = w0, w4 = w2, w5 = w3
        w[2] = np.swapaxes(acc[0], 0, 1)
        w[1], w[4], w[5] = w[0], w[2], w[3]
    elif mode == "ijj_acc":             # w3 = w2(a,c,b); then w1 = w0, w4 = w2, w5 = w3
        w[3] = np.swapaxes(acc[2], 1, 2)
        w[1], w[4], w[5] = w[0], w[2], w[3]
    elif mode == "iij_aac":             # w4 = w1(b,a,c); then w2 = w0, w3 = w1, w5 = w4
        w[4] = np.swapaxes(acc[1], 0, 1)
        w[2], w[3], w[5] = w[0], w[1], w[4]
    elif mode == "iij_acc":             # w1 = w0(a,c,b); then w2 = w0, w3 = w1, w5 = w4
        w[1] = np.swapaxes(acc[0], 1, 2)
        w[2], w[3], w[5] = w[0], w[1], w[4]
    elif mode == "aac":                 # w2(a,b,c)=w0(b,a,c), w4=w1(b,a,c), w5=w3(b,a,c)
        w[2], w[4], w[5] = (np.swapaxes(acc[q], 0, 1) for q in (0, 1, 3))
    elif mode == "acc":                 # w1=w0(a,c,b), w3=w2(a,c,b), w5=w4(a,c,b)
        w[1], w[3], w[5] = (np.swapaxes(acc[q], 1, 2) for q in (0, 2, 4))
    elif mode == "aaa":                 # w_P(e) = w_0(P^-1 . e)
        for p in range(1, 6):
            inv = [PERM[p].index(m) for m in range(3)]
            # out[e0,e1,e2] = acc0[e_inv0, e_inv1, e_inv2]
            w[p] = np.einsum("xyz->" + "".join("xyz"[inv.index(n)] for n in range(3)), acc[0])
    return w


def blocks(nb: int) -> list:
    out = []
    for a in range(nb):
        for b in range(a, nb):
            for c in range(b, nb):
                m = 6 if a == b == c else 2 if (a == b or b == c) else 1
                out.append((a, b, c, m))
    return out


def model(inp: dict, tb: int = 4, pair: bool = True) -> dict:
    no, nu = inp["no"], inp["nu"]
    nup = -(-nu // tb) * tb
    nb = nup // tb
    kk = nu + no
    vvvo, t2, vm, voe = inp["vvvo"], inp["t2"], inp["vm"], inp["voe"]
    ax = np.zeros((no, nup, nup, kk))
    ax[:, :nu, :nu, :nu] = np.einsum("kbax->xabk", vvvo)          # AX[x](a,b,K) = VVVO(K,b,a,x)
    ax[:, :nu, :nu, nu:] = np.einsum("ablx->xabl", t2)            # AX[x](a,b,nu+l) = T2(a,b,l,x)
    bx = np.zeros((no, no, nup, kk))
    bx[:, :, :nu, :nu] = np.einsum("gkyz->yzgk", t2)              # BX[y,z](g,K) = T2(g,K,y,z)
    bx[:, :, :nu, nu:] = -np.einsum("lgzy->yzgl", vm)             # BX[y,z](g,nu+l) = -VM(l,g,z,y)
    voep = np.zeros((no, no, nup, nup))
    voep[:, :, :nu, :nu] = np.einsum("xyqr->qrxy", voe)
    ep = np.concatenate([inp["ep"], np.full(nup - nu, 1e3)])
    eh = inp["eh"]
    etd = 0.0
    v1 = np.zeros((nu, no))
    for (ti, tj, tk, kind) in tuples(no):
        tp = (ti, tj, tk)
        for (A, B, C, mult) in blocks(nb):
            A, B, C = ragged_to_b(nu, tb, A, B, C, mult)
            sa, sb, sc = (slice(X * tb, X * tb + tb) for X in (A, B, C))
            acc = np.zeros((6, tb, tb, tb))
            mode = item_mode(kind, mult, A, B, pair)
            for p in MODE_SET[mode]:
                for phase in range(3):
                    s1, b1, s2, b2 = term(phase, p)
                    for (s, bq) in ((s1, b1), (s2, b2)):
                        x = tp[s % 3]
                        o = s // 3
                        y, z = [(yy, zz) for yy in range(3) for zz in range(3)
                                if yy != zz and pair_index(yy, zz) == bq][0]
                        bxm = bx[tp[y], tp[z]]
                        if phase == 0:      # gamma=b
                            pan = ax[x][sa, sc] if o == 0 else np.swapaxes(ax[x][sc, sa], 0, 1)
                            acc[p] += np.einsum("ack,bk->abc", pan, bxm[sb])
                        elif phase == 1:    # gamma=c
                            pan = ax[x][sa, sb] if o == 0 else np.swapaxes(ax[x][sb, sa], 0, 1)
                            acc[p] += np.einsum("abk,ck->abc", pan, bxm[sc])
                        else:               # gamma=a
                            pan = ax[x][sb, sc] if o == 0 else np.swapaxes(ax[x][sc, sb], 0, 1)
                            acc[p] += np.einsum("bck,ak->abc", pan, bxm[sa])
            acc = complete(acc, mode)
            wgt = (1.0 if kind == 1 else 0.5) / mult
            dijk = eh[ti] + eh[tj] + eh[tk]
            for al in range(tb):
                for bl in range(tb):
                    for cl in range(tb):
                        ag, bg, cg = A * tb + al, B * tb + bl, C * tb + cl
                        if ag >= nu or bg >= nu or cg >= nu:
                            continue
                        w = acc[:, al, bl, cl]
                        inv = 1.0 / (dijk - ep[ag] - ep[bg] - ep[cg])
                        ev, od = w[0] + w[3] + w[4], w[1] + w[2] + w[5]
                        q = 6.0 * np.dot(w, w) + 2.0 * ev * ev + 2.0 * od * od - 8.0 * ev * od
                        etd += q * inv * wgt
                        el = (ag, bg, cg)
                        for s in range(6):
                            r0, r1, r2 = PERM[s]
                            for p in range(6):
                                zv = (2.0 * w[wsrc_comp(p, s, 0)] - 2.0 * w[wsrc_comp(p, s, 2)]
                                      - w[wsrc_comp(p, s, 1)] + w[wsrc_comp(p, s, 3)])
                                vv = voep[tp[PERM[p][1]], tp[PERM[p][2]], el[r1], el[r2]]
                                v1[el[r0], tp[PERM[p][0]]] += zv * vv * inv * wgt
    ets = 2.0 * float(np.einsum("pa,ap->", inp["t1"], v1))
    return {"etd": etd, "ets": ets, "v1": v1}
