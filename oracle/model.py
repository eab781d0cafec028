"""Oracle model parameterisations (test infrastructure only).

Restates ADFWI/model/parameters.py:47-213 and the Gardner density relation of
ADFWI/model/acoustic_model.py:111-118, so that coefficient fields fed to both
the oracle propagators and the product can be built from vp/vs/rho/eps/delta.
"""
import numpy as np
import torch


def gardner_rho(vp: np.ndarray) -> np.ndarray:
    """rho = 310 * vp^0.25 (acoustic_model.py:115), float32 numpy like the reference."""
    return np.power(vp, 0.25) * 310


def lame(vp, vs, rho):
    """parameters.py:47-68."""
    mu = vs ** 2 * rho
    lamu = vp ** 2 * rho
    lam = lamu - 2 * mu
    b = 1 / rho
    return mu, lamu, lam, b


def thomsen_to_cij(vp, vs, rho, eps, delta, gamma):
    """parameters.py:71-107."""
    c33 = vp ** 2 * rho
    c44 = vs ** 2 * rho
    c11 = c33 * (1 + 2 * eps)
    c66 = c44 * (1 + 2 * gamma)
    c13 = torch.sqrt(2 * c33 * (c33 - c44) * delta + (c33 - c44) ** 2) - c44
    return c11, c13, c33, c44, c66


def staggered(mu, b, c44, c55, c66, nx, nz):
    """parameters.py:184-213: bx, bz two-point averages and the 5-term shear averages."""
    bx = 0.5 * (b[:, 0:nx - 1] + b[:, 1:nx])
    bz = 0.5 * (b[0:nz - 1, :] + b[1:nz, :])

    def five(m):
        return 0.2 * (m[1:nz - 1, 1:nx - 1] + m[2:nz, 1:nx - 1] + m[1:nz - 1, 2:nx]
                      + m[2:nz, 1:nx - 1] + m[2:nz, 2:nx])
    return bx, bz, five(mu), five(c44), five(c55), five(c66)


def elastic_cc(vp, vs, rho, eps, delta, gamma, nx, nz):
    """Isotropic/VTI CC list of 21 moduli as IsotropicElasticModel / AnisotropicElasticModel
    build it (elastic_model.py:147-164, 392-407; parameters.py:135-181, VTI branch).

    Returns (CC, bx, bz, lamu, lam).
    """
    mu, lamu, lam, b = lame(vp, vs, rho)
    z = torch.zeros((nz, nx), dtype=vp.dtype, device=vp.device)
    c11, c13, c33, c44, c66 = thomsen_to_cij(vp, vs, rho, eps, delta, gamma)
    c55 = c44.clone()
    bx, bz, _, c44s, c55s, c66s = staggered(mu, b, c44, c55, c66, nx, nz)
    c22 = c11.clone()
    c23 = c13.clone()
    c12 = (c11 - 2 * c66).clone()
    CC = [c11, c12, c13, z, z, z, c22, c23, z, z, z, c33, z, z, z, c44s, z, z, c55s, z, c66s]
    return CC, bx, bz, lamu, lam
