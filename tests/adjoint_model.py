"""numpy model of the per-tuple adjoint the GPU path computes (test helper).

For each tuple t (canonical order, weight 1 or 1/2): W_t from the closed form,
Wbar_t = wt*(dX/dW + dE_ST/dW) elementwise (the tables of ccx_adjoint.hip), dVOE from Z,
then the six sigma-terms of the closed form back-contracted into the AX/BX slab
gradients, finally unpacked into dVVVO, dT2, dVM.  dT1 = 2 v1^T."""

from __future__ import annotations

import numpy as np

PERM = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
# Z = (2 W(x) - 2 W(t12 x) - W(t23 x) + W(rho x)) / D, (mu x)_m = x_mu[m]
ZTERMS = [((0, 1, 2), 2.0), ((1, 0, 2), -2.0), ((0, 2, 1), -1.0), ((1, 2, 0), 1.0)]
# sigma-terms of W: (x position, AX index positions (u, v), BX pair positions (y, z), BX free position w)
SIGMA = [(0, (0, 1), (1, 2), 2), (1, (1, 0), (0, 2), 2), (2, (2, 1), (1, 0), 0),
         (0, (0, 2), (2, 1), 1), (1, (1, 2), (2, 0), 0), (2, (2, 0), (0, 1), 1)]


def inv(p):
    r = [0, 0, 0]
    for m in range(3):
        r[p[m]] = m
    return r


def tuples(no):
    out = [((i, j, k), 1.0) for i in range(no) for j in range(i) for k in range(j)]
    for i in range(no):
        for j in range(i):
            out.append(((i, j, j), 0.5))
            out.append(((i, i, j), 0.5))
    return out


def w_tuple(inp, t):
    i, j, k = t
    vvvo, t2, vm = inp["vvvo"], inp["t2"], inp["vm"]
    nu = inp["nu"]
    ax = {x: np.concatenate([np.einsum("kba->abk", vvvo[:, :, :, x]), t2[:, :, :, x]], axis=2) for x in set(t)}
    bx = {(y, z): np.concatenate([t2[:, :, y, z], -vm[:, :, z, y].T], axis=1) for y in set(t) for z in set(t)}
    w = np.zeros((nu, nu, nu))
    letters = "abc"
    for (xp, (u, v), (yp, zp), wf) in SIGMA:
        s = "".join(letters[q] for q in (u, v)) + "K," + letters[wf] + "K->abc"
        w += np.einsum(s, ax[t[xp]], bx[(t[yp], t[zp])])
    return w, ax, bx


def model_grad(inp):
    no, nu = inp["no"], inp["nu"]
    eh, ep, t1, voe = inp["eh"], inp["ep"], inp["t1"], inp["voe"]
    kk = nu + no
    g_ax = np.zeros((no, nu, nu, kk))
    g_bx = np.zeros((no, no, nu, kk))
    d_voe = np.zeros_like(voe)
    a_idx = np.arange(nu)
    dv = ep[:, None, None] + ep[None, :, None] + ep[None, None, :]
    for t, wt in tuples(no):
        w, ax, bx = w_tuple(inp, t)
        d = eh[t[0]] + eh[t[1]] + eh[t[2]] - dv

        def wp(s):
            return np.einsum(s + "->abc", w)
        dx = (16 * w - 8 * (wp("acb") + wp("cba") + wp("bac")) + 4 * (wp("bca") + wp("cab"))) / d
        dx[a_idx, a_idx, a_idx] = 0.0
        wbar = dx.copy()
        # E_ST adjoint: sum over position perms pi (ordering P = pi.t) and Z terms mu
        for pi in PERM:
            P = tuple(t[pi[m]] for m in range(3))
            pinv = inv(pi)
            for (mu, cm) in ZTERMS:
                lam = [mu[pinv[n]] for n in range(3)]          # n -> mu[pi^-1[n]]
                li = inv(lam)
                # Y_P(x*) with x*_j = e_{lam^-1[j]} ; Y_P(x) = 2 T1(P0,x0) VOE(x1,x2,P1,P2)
                # build as a cube over e = (a,b,c)
                letters = "abc"
                x0, x1, x2 = (letters[li[j]] for j in range(3))
                y = 2.0 * np.einsum(f"{x0},{x1}{x2}->abc", t1[P[0]], voe[:, :, P[1], P[2]])
                wbar += cm * y / d
            # dVOE(x1,x2,P1,P2) += 2 wt sum_x0 T1(P0,x0) Z_P(x)
            zp = np.zeros((nu, nu, nu))
            for (mu, cm) in ZTERMS:
                lam = [mu[pinv[n]] for n in range(3)]
                # W_P(mu x) = W_t(pi^-1 mu x): (pi^-1 mu x)_n = x_{mu[pi^-1[n]]} = x_{lam[n]}
                zp += cm * _gather(w, lam)
            zp /= d
            d_voe[:, :, P[1], P[2]] += 2.0 * wt * np.einsum("a,abc->bc", t1[P[0]], zp)
        wbar *= wt
        for (xp, (u, v), (yp, zp_), wf) in SIGMA:
            letters = "abc"
            uv = letters[u] + letters[v]
            g_ax[t[xp]] += np.einsum(f"abc,{letters[wf]}K->{uv}K", wbar, bx[(t[yp], t[zp_])])
            g_bx[t[yp], t[zp_]] += np.einsum(f"abc,{uv}K->{letters[wf]}K", wbar, ax[t[xp]])
    dvvvo = np.einsum("xabk->kbax", g_ax[:, :, :, :nu])
    dt2 = np.einsum("xabl->ablx", g_ax[:, :, :, nu:]) + np.einsum("yzgk->gkyz", g_bx[:, :, :, :nu])
    dvm = -np.einsum("yzgl->lgzy", g_bx[:, :, :, nu:])
    return {"vvvo": dvvvo, "t2": dt2, "vm": dvm, "voe": d_voe}


def _gather(w, lam):
    """out[x0,x1,x2] = w[x_lam[0], x_lam[1], x_lam[2]]"""
    s_in = "".join("xyz"[lam[n]] for n in range(3))
    return np.einsum(s_in + "->xyz", w)
