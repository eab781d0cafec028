"""Absorbing-boundary profiles (host numpy, once per propagator).

Vectorised equivalents of ADFWI/propagator/boundary_condition.py:15-120 with
the same float64 operation order, so the profiles agree bit for bit.
"""
import math

import numpy as np


def _dims(nx, nz, pml, free_surface):
    return nx + 2 * pml, (nz + pml if free_surface else nz + 2 * pml)


def bc_pml(nx, nz, dx, dz, pml, vmax, free_surface=True):
    """Quadratic PML damping d = kappa (k dx / a)^2 with a = (pml-1) dx, R = 1e-6
    (boundary_condition.py:15-47).  The x strips are laid first and the z
    trapezoids overwrite them (top one only without a free surface)."""
    nxp, nzp = _dims(nx, nz, pml, free_surface)
    a = (pml - 1) * dx
    kappa = -3.0 * vmax * math.log(1e-6) / (2.0 * a)
    xa = np.arange(pml) * dx / a
    prof = kappa * xa * xa
    d = np.zeros((nzp, nxp))
    k = np.arange(pml)
    d[:, pml - 1 - k] = prof[k]
    d[:, nxp - pml + k] = prof[k]
    # trapezoid row of depth k covers columns [pml-k, nxp-pml+k)
    cols = np.arange(nxp)
    inside = (cols[None, :] >= (pml - k)[:, None]) & (cols[None, :] < (nxp - pml + k)[:, None])
    vals = np.where(inside, prof[:, None], 0.0)
    bands = [nzp - pml + k]                 # bottom rows, depth k
    if not free_surface:
        bands.append(pml - 1 - k)           # top rows, depth k
    for rows in bands:
        d[rows, :] = np.where(inside, vals, d[rows, :])
    return d


def bc_pml_xz(nx, nz, dx, dz, pml, vmax, free_surface=True):
    """Split-field profiles BCx, BCz = (pml-k+1)^2 ppml / dx, ppml = -ln(1e-6) 3 vmax / (2 pml^3)
    (boundary_condition.py:98-120); no top z strip with a free surface."""
    nxp, nzp = _dims(nx, nz, pml, free_surface)
    ppml = -np.log(1e-6) * 3 * vmax / (2 * pml ** 3)
    k = np.arange(1, pml + 1)
    bcx = np.zeros((nzp, nxp))
    bcz = np.zeros((nzp, nxp))
    vx = (pml - k + 1) ** 2 * ppml / dx
    vz = (pml - k + 1) ** 2 * ppml / dz
    bcx[:, k - 1] = vx
    bcx[:, nxp - k] = vx
    if not free_surface:
        bcz[k - 1, :] = vz[:, None]
    bcz[nzp - k, :] = vz[:, None]
    return bcx, bcz


def bc_sincos(nx, nz, dx, dz, pml, free_surface=False):
    """Multiplicative sin^2 taper (boundary_condition.py:50-68)."""
    nxp, nzp = _dims(nx, nz, pml, free_surface)
    d = np.ones((nzp, nxp))
    for i in range(pml):
        f = np.sin(np.pi / 2 * i / pml) ** 2
        if not free_surface:
            d[i, :] *= f
        d[-i - 1, :] *= f
        d[:, i] *= f
        d[:, -i - 1] *= f
    return d


def bc_gerjan(nx, nz, dx, dz, pml, alpha=0.0053, free_surface=True):
    """Cerjan taper exp(-(alpha (pml - k))^2) on nested frames (boundary_condition.py:71-95)."""
    wt = np.exp(-(alpha * (pml - np.arange(1, pml + 1))) ** 2)
    nxp, nzp = _dims(nx, nz, pml, free_surface)
    d = np.ones((nzp, nxp))
    for k in range(1, pml + 1):
        w = wt[k - 1]
        if free_surface:
            d[:nzp - k + 1, k - 1] = w
            d[nzp - k, k - 1:nxp - k + 1] = w
            d[:nzp - k + 1, nxp - k] = w
        else:
            d[k - 1:nzp - k + 1, k - 1] = w
            d[k:nzp - k + 1, nxp - k] = w
            d[nzp - k, k - 1:nxp - k + 1] = w
            d[k - 1, k - 1:nxp - k + 1] = w
    return d
