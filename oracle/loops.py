"""Pure-Python loop restatement of one acoustic time step, for tiny grids only.

Test infrastructure: an independent formulation (explicit index loops instead
of slices) of ADFWI/propagator/acoustic_kernels.py:115-169, used to pin the
window bounds and free-surface rows of the vectorised oracle and of the HIP
kernel.  float64 numpy.
"""
import numpy as np

C1 = 9.0 / 8.0
C2 = -1.0 / 24.0


def acoustic_step(p, u, w, omk1, a1, omk2, a2, omk3, fs, free_surface, src, rcv):
    """One step for one shot.  p (Z, X), u (Z, X-1), w (Z-1, X) are updated copies.

    src: (zs, xs, value) injected after the pressure update; rcv: list of (z, x).
    Returns p, u, w and the records [(p, u, w) at each receiver].
    """
    p, u, w = p.copy(), u.copy(), w.copy()
    Z, X = p.shape
    pn = p.copy()
    for i in range(Z):
        for j in range(X):
            if fs + 1 <= i <= Z - 3 and 2 <= j <= X - 3:
                d = C1 * (u[i, j] - u[i, j - 1] + w[i, j] - w[i - 1, j]) + \
                    C2 * (u[i, j + 1] - u[i, j - 2] + w[i + 1, j] - w[i - 2, j])
                pn[i, j] = omk1[i, j] * p[i, j] - a1[i, j] * d
    zs, xs, val = src
    pn[zs, xs] += val
    if free_surface:
        for j in range(X):
            pn[fs - 1, j] = -pn[fs + 1, j]
    un = u.copy()
    for i in range(Z):
        for j in range(X - 1):
            if fs <= i <= Z - 2 and 1 <= j <= X - 3:
                un[i, j] = omk2[i, j] * u[i, j] - a2[i, j] * (C1 * (pn[i, j + 1] - pn[i, j]) +
                                                             C2 * (pn[i, j + 2] - pn[i, j - 1]))
    wn = w.copy()
    for i in range(Z - 1):
        for j in range(X):
            if fs <= i <= Z - 3 and 1 <= j <= X - 2:
                wn[i, j] = omk3[i, j] * w[i, j] - a2[i, j] * (C1 * (pn[i + 1, j] - pn[i, j]) +
                                                             C2 * (pn[i + 2, j] - pn[i - 1, j]))
    if free_surface:
        for j in range(X):
            wn[fs - 1, j] = wn[fs, j]
    recs = [(pn[z, x], un[z, x], wn[z, x]) for z, x in rcv]
    return pn, un, wn, recs
