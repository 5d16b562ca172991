"""NumPy restatement of the learned initial guess, for checking the HIP
kernels (test infrastructure only).  Float64 throughout.

Model: x0_i = exp(i kappa d_i) (b_i + exp(i k z_i) NN(f_i)); NN is an MLP with
`nlayers` SiLU layers of width `hidden` and a linear output; parameters are a
flat array, per layer W [out x in] row-major then b [out], then kappa (re, im).
"""
from __future__ import annotations

import numpy as np

NFEAT, NSHELL, SHELLF, GLOBALF = 64, 4, 14, 6


def amn(m, n, L):
    """MSTM model-2 position of (m, n) inside a mode block (mstm-modules-33.f90:2365-2377), 0-based."""
    return (n - 1) * (L + 2) + m if m >= 0 else (-m - 1) * (L + 2) + n + 1


def nparams(hidden, nlayers, nout, nin=NFEAT):
    n, i = 0, nin
    for l in range(nlayers + 1):
        o = nout if l == nlayers else hidden
        n += o * i + o
        i = o
    return n + 2


def unpack(p, hidden, nlayers, nout, nin=NFEAT):
    Ws, bs, off, i = [], [], 0, nin
    for l in range(nlayers + 1):
        o = nout if l == nlayers else hidden
        Ws.append(np.asarray(p[off:off + o * i], np.float64).reshape(o, i))
        off += o * i
        bs.append(np.asarray(p[off:off + o], np.float64))
        off += o
        i = o
    return Ws, bs, np.asarray(p[off:off + 2], np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, hidden, nlayers, nout):
    Ws, bs, _ = unpack(p, hidden, nlayers, nout)
    h = np.asarray(X, np.float64)
    Zs, Hs = [], [h]
    for l in range(nlayers + 1):
        z = h @ Ws[l].T + bs[l]
        if l == nlayers:
            return z, Zs, Hs
        Zs.append(z)
        h = silu(z)
        Hs.append(h)


def loss_grad(p, X, Y, base, depth, hidden, nlayers, nout):
    """Mean over samples of |yhat - y|^2 / mean|y|^2 and its gradient."""
    Ws, bs, kap = unpack(p, hidden, nlayers, nout)
    O, Zs, Hs = forward(p, X, hidden, nlayers, nout)
    B = len(X)
    yscale = float(np.sum(np.asarray(Y, np.float64) ** 2)) / B
    norm = 1.0 / (B * yscale)
    Oc = O[:, 0::2] + 1j * O[:, 1::2]
    Yc = np.asarray(Y, np.float64)[:, 0::2] + 1j * np.asarray(Y, np.float64)[:, 1::2]
    Bc = np.asarray(base, np.float64)[:, 0::2] + 1j * np.asarray(base, np.float64)[:, 1::2]
    d = np.asarray(depth, np.float64)[:, None]
    E = np.exp(1j * (kap[0] + 1j * kap[1]) * d)
    yhat = E * (Bc + Oc)
    e = yhat - Yc
    loss = float(np.sum(np.abs(e) ** 2)) * norm
    g = np.conj(e) * E
    dO = np.zeros_like(O)
    dO[:, 0::2] = 2.0 * g.real * norm
    dO[:, 1::2] = -2.0 * g.imag * norm
    h = np.conj(e) * yhat
    dk = np.array([np.sum(-2.0 * d * h.imag), np.sum(-2.0 * d * h.real)]) * norm
    grads = []
    dA = dO
    for l in range(nlayers, -1, -1):
        gW = dA.T @ Hs[l]
        gb = dA.sum(0)
        grads.append((gW, gb))
        if l > 0:
            dA = (dA @ Ws[l]) * silu_grad(Zs[l - 1])
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    flat.append(dk)
    return loss, np.concatenate(flat)


def features(pos, radius, ri, rimed=1.0):
    """Per-sphere features and depths, the same definitions as lig_features_kernel."""
    pos = np.asarray(pos, np.float64)
    radius = np.asarray(radius, np.float64)
    ri = np.asarray(ri, np.complex128)
    n = len(radius)
    F = np.zeros((n, NFEAT))
    depth = np.zeros(n)
    k = rimed
    for i in range(n):
        dv = pos - pos[i]
        ai = radius[i]
        D, rc = 2 * ai, 8 * ai
        rho2 = dv[:, 0] ** 2 + dv[:, 1] ** 2
        col = rho2 < (ai + radius) ** 2
        zmin = min(pos[i, 2] - ai, np.min(pos[col, 2] - radius[col]))
        zmax = max(pos[i, 2] + ai, np.max(pos[col, 2] + radius[col]))
        r2 = rho2 + dv[:, 2] ** 2
        m = (r2 < rc * rc)
        m[i] = False
        edges = [0.0, 1.25 * D, 2.0 * D, 3.0 * D, rc]
        for s in range(NSHELL):
            sel = m & (r2 >= edges[s] ** 2) & (r2 < edges[s + 1] ** 2)
            r = np.sqrt(r2[sel])
            u = dv[sel] / r[:, None]
            ux, uy, uz = u[:, 0], u[:, 1], u[:, 2]
            dz = dv[sel, 2]
            wr = D / r
            acc = [len(r), ux.sum(), uy.sum(), uz.sum(), (ux * ux - uy * uy).sum(),
                   (2 * uz * uz - ux * ux - uy * uy).sum(), (ux * uy).sum(), (ux * uz).sum(), (uy * uz).sum(),
                   np.cos(k * dz).sum(), np.sin(k * dz).sum(), (wr * np.cos(k * (dz + r))).sum(),
                   (wr * np.sin(k * (dz + r))).sum(), (wr ** 3).sum()]
            F[i, GLOBALF + s * SHELLF:GLOBALF + (s + 1) * SHELLF] = np.array(acc) / 20.0
        F[i, 0] = min((pos[i, 2] - zmin) / rc, 1.0)
        F[i, 1] = min((zmax - pos[i, 2]) / rc, 1.0)
        F[i, 2] = k * ai / 4.0
        F[i, 3] = ri[i].real / k - 1.0
        F[i, 4] = ri[i].imag / k
        F[i, 5] = m.sum() / 150.0
        depth[i] = pos[i, 2] - zmin
    return F, depth


def to_rows(vec, pos, order, offset, lnet, rimed=1.0):
    """Per-sphere rows (8 lnet(lnet+2) floats) of vec(2, neqns) * exp(-i k z_i)."""
    NBn = lnet * (lnet + 2)
    n = len(order)
    out = np.zeros((n, 8 * NBn))
    for i in range(n):
        L = order[i]
        NB = L * (L + 2)
        e = np.exp(-1j * rimed * pos[i, 2])
        for p in range(2):
            for mode in range(2):
                for nn in range(1, min(L, lnet) + 1):
                    for m in range(-nn, nn + 1):
                        v = vec[p, offset[i] + mode * NB + amn(m, nn, L)] * e
                        dst = p * 4 * NBn + 2 * (mode * NBn + amn(m, nn, lnet))
                        out[i, dst], out[i, dst + 1] = v.real, v.imag
    return out


def combine(b, O, depth, kappa, pos, order, offset, lnet, rimed=1.0):
    """x0(2, neqns) from b(2, neqns) and network outputs O (n rows)."""
    NBn = lnet * (lnet + 2)
    x = np.array(b, np.complex128)
    for i in range(len(order)):
        L = order[i]
        NB = L * (L + 2)
        eph = np.exp(1j * rimed * pos[i, 2])
        E = np.exp(1j * (kappa[0] + 1j * kappa[1]) * depth[i])
        for p in range(2):
            for mode in range(2):
                for nn in range(1, L + 1):
                    for m in range(-nn, nn + 1):
                        idx = offset[i] + mode * NB + amn(m, nn, L)
                        v = b[p, idx]
                        if nn <= lnet:
                            s = p * 4 * NBn + 2 * (mode * NBn + amn(m, nn, lnet))
                            v = v + eph * (O[i, s] + 1j * O[i, s + 1])
                        x[p, idx] = E * v
    return x
