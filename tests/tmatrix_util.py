"""Helpers for T-matrix tests: the layout of tmatrix_solution's file body
(mstm-solver-8.f90:243-256) and the reciprocity check."""
import numpy as np


def tmat_size(T):
    return sum(4 * (2 * l + 1) * l * (l + 2) for l in range(1, T + 1))


def unpack(tmat, T):
    """{(n, m, p, l, k, q): value}: columns in write order (l, k = -l..l,
    q = 1, 2), rows n = 1..l, m = -n..n, p = 1, 2."""
    out, i = {}, 0
    for l in range(1, T + 1):
        for k in range(-l, l + 1):
            for q in (1, 2):
                for n in range(1, l + 1):
                    for m in range(-n, n + 1):
                        for p in (1, 2):
                            out[(n, m, p, l, k, q)] = tmat[i]
                            i += 1
    assert i == len(tmat)
    return out


def reciprocity_error(tmat, T):
    """max |T^{pq}_{mn,kl} - (-1)^{m+k} T^{qp}_{-k l,-m n}| / max |T| over the
    entries whose reciprocal partner is stored (the blocks n = l)."""
    t = unpack(tmat, T)
    err = 0.0
    for (n, m, p, l, k, q), v in t.items():
        w = t.get((l, -k, q, n, -m, p))
        if w is not None:
            err = max(err, abs(v - (-1) ** (m + k) * w))
    return err / float(np.abs(tmat).max())
