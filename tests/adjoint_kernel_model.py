"""numpy model of the adjoint work-item epilogue (test helper).

It repeats, with loops over (tuple, block-orbit) items, the tables that
libcc_amd/csrc/ccx_adjoint.hip uses: per point of an item the six W values give
Wbar at the six positional permutations of the point, and the 36 (placement,
ordering) products give the T1.VOE part of Wbar and the dVOE contributions.  The
Wbar cubes then feed the sigma-term back-contractions.  Comparing with
adjoint_model.model_grad checks the tables on the CPU."""

from __future__ import annotations

import numpy as np

from kernel_model import (PERM, MODE_SET, blocks, complete, item_mode, pair_index, perm_index, ragged_to_b, term,
                          tuples, wsrc_comp)
from adjoint_model import SIGMA, ZTERMS, inv


def sn_table():
    """sn[s][p][m]: placement index n = PERM[s] o li of the product for Z term m of ordering p."""
    out = np.zeros((6, 6, 4), dtype=int)
    for s in range(6):
        for p in range(6):
            pinv = inv(PERM[p])
            for m, (mu, _) in enumerate(ZTERMS):
                lam = [mu[pinv[n]] for n in range(3)]
                li = inv(lam)
                n = [PERM[s][li[k]] for k in range(3)]
                out[s, p, m] = perm_index(*n)
    return out


def model_adjoint(inp: dict, tb: int = 4) -> dict:
    no, nu = inp["no"], inp["nu"]
    nup = -(-nu // tb) * tb
    nb = nup // tb
    kk = nu + no
    vvvo, t2, vm, voe, t1 = inp["vvvo"], inp["t2"], inp["vm"], inp["voe"], inp["t1"]
    ax = np.zeros((no, nup, nup, kk))
    ax[:, :nu, :nu, :nu] = np.einsum("kbax->xabk", vvvo)
    ax[:, :nu, :nu, nu:] = np.einsum("ablx->xabl", t2)
    bx = np.zeros((no, no, nup, kk))
    bx[:, :, :nu, :nu] = np.einsum("gkyz->yzgk", t2)
    bx[:, :, :nu, nu:] = -np.einsum("lgzy->yzgl", vm)
    voep = np.zeros((nup, nup, no, no))
    voep[:nu, :nu] = voe
    t1p = np.zeros((no, nup))
    t1p[:, :nu] = t1
    ep = np.concatenate([inp["ep"], np.full(nup - nu, 1e3)])
    eh = inp["eh"]
    sn = sn_table()
    cmu = [c for (_, c) in ZTERMS]
    g_ax = np.zeros((no, nup, nup, kk))
    g_bx = np.zeros((no, no, nup, kk))
    d_voe = np.zeros((nup, nup, no, no))
    for (ti, tj, tk, kind) in tuples(no):
        tp = (ti, tj, tk)
        wt = 1.0 if kind == 1 else 0.5
        wbar = np.zeros((nup, nup, nup))
        for (A, B, C, mult) in blocks(nb):
            blk = ragged_to_b(nu, tb, A, B, C, mult)
            acc = np.zeros((6, tb, tb, tb))
            mode = item_mode(kind, mult, blk[0], blk[1])
            for p in MODE_SET[mode]:
                for phase in range(3):
                    s1, b1, s2, b2 = term(phase, p)
                    for (s, bq) in ((s1, b1), (s2, b2)):
                        x = tp[s % 3]
                        o = s // 3
                        y, z = [(yy, zz) for yy in range(3) for zz in range(3)
                                if yy != zz and pair_index(yy, zz) == bq][0]
                        bxm = bx[tp[y], tp[z]]
                        sa, sb, sc = (slice(X * tb, X * tb + tb) for X in blk)
                        if phase == 0:
                            pan = ax[x][sa, sc] if o == 0 else np.swapaxes(ax[x][sc, sa], 0, 1)
                            acc[p] += np.einsum("ack,bk->abc", pan, bxm[sb])
                        elif phase == 1:
                            pan = ax[x][sa, sb] if o == 0 else np.swapaxes(ax[x][sb, sa], 0, 1)
                            acc[p] += np.einsum("abk,ck->abc", pan, bxm[sc])
                        else:
                            pan = ax[x][sb, sc] if o == 0 else np.swapaxes(ax[x][sc, sb], 0, 1)
                            acc[p] += np.einsum("bck,ak->abc", pan, bxm[sa])
            acc = complete(acc, mode)
            wgt = wt / mult
            dijk = eh[ti] + eh[tj] + eh[tk]
            for al in range(tb):
                for bl in range(tb):
                    for cl in range(tb):
                        el = (al, bl, cl)
                        xg = tuple(blk[n] * tb + el[n] for n in range(3))
                        w = acc[:, al, bl, cl]
                        invd = 1.0 / (dijk - ep[xg[0]] - ep[xg[1]] - ep[xg[2]])
                        wt_s = [w[wsrc_comp(0, s, 0)] for s in range(6)]
                        ev, od = w[0] + w[3] + w[4], w[1] + w[2] + w[5]
                        t1v = [[t1p[tp[y], xg[n]] for n in range(3)] for y in range(3)]
                        prod = np.zeros((6, 6))
                        for s in range(6):
                            r = PERM[s]
                            for p in range(6):
                                P = PERM[p]
                                tv = t1v[P[0]][r[0]]
                                vv = voep[xg[r[1]], xg[r[2]], tp[P[1]], tp[P[2]]]
                                prod[s, p] = tv * vv
                                zv = (2.0 * w[wsrc_comp(p, s, 0)] - 2.0 * w[wsrc_comp(p, s, 2)]
                                      - w[wsrc_comp(p, s, 1)] + w[wsrc_comp(p, s, 3)])
                                d_voe[xg[r[1]], xg[r[2]], tp[P[1]], tp[P[2]]] += 2.0 * wgt * zv * invd * tv
                        diag = xg[0] == xg[1] == xg[2]
                        for s in range(6):
                            even = s in (0, 3, 4)
                            same, other = (ev, od) if even else (od, ev)
                            dx = 0.0 if diag else (12.0 * wt_s[s] + 4.0 * same - 8.0 * other) * invd
                            ets = 0.0
                            for p in range(6):
                                for m in range(4):
                                    ets += cmu[m] * prod[sn[s, p, m], p]
                            xs = tuple(xg[PERM[s][m]] for m in range(3))
                            wbar[xs] = 0.0 if diag else wt * (dx + 2.0 * invd * ets)
        for (xp, (u, v), (yp, zp), wf) in SIGMA:
            letters = "abc"
            uv = letters[u] + letters[v]
            g_ax[tp[xp]] += np.einsum(f"abc,{letters[wf]}K->{uv}K", wbar, bx[tp[yp], tp[zp]])
            g_bx[tp[yp], tp[zp]] += np.einsum(f"abc,{uv}K->{letters[wf]}K", wbar, ax[tp[xp]])
    g_ax, g_bx = g_ax[:, :nu, :nu], g_bx[:, :, :nu]
    dvvvo = np.einsum("xabk->kbax", g_ax[:, :, :, :nu])
    dt2 = np.einsum("xabl->ablx", g_ax[:, :, :, nu:]) + np.einsum("yzgk->gkyz", g_bx[:, :, :, :nu])
    dvm = -np.einsum("yzgl->lgzy", g_bx[:, :, :, nu:])
    return {"vvvo": dvvvo, "t2": dt2, "vm": dvm, "voe": d_voe[:nu, :nu]}
