"""Elastic parameterisations that produce the propagator's coefficient fields.

Same functions and formulas as ADFWI/model/parameters.py:12-213 (they define
the inputs of the hot path, so they must agree to the bit): Lame constants,
Thomsen -> stiffness, isotropic / TI filling of the 21 moduli, staggered-grid
averages of buoyancy and shear moduli.
"""
import torch

from ..utils import numpy2tensor


def thomsen_init(vp, vs, rho, eps, delta, gamma, device, dtype=torch.float32):
    return tuple(numpy2tensor(a, dtype).to(device) for a in (vp, vs, rho, eps, delta, gamma))


def elastic_moduli_init(nz, nx, device, dtype=torch.float32):
    """21 zero (nz, nx) moduli in the order C11..C16, C22..C26, C33..C36, C44..C46, C55, C56, C66."""
    return [torch.zeros((nz, nx), dtype=dtype).to(device) for _ in range(21)]


def vs_vp_to_Lame(vp, vs, rho):
    """mu = vs^2 rho, lambda+2mu = vp^2 rho, lambda, buoyancy 1/rho (parameters.py:47-68)."""
    mu = vs ** 2 * rho
    lamu = vp ** 2 * rho
    lam = lamu - 2 * mu
    b = 1 / rho
    return mu, lamu, lam, b


def thomsen_to_elastic_moduli(alpha_thomsen, beta_thomsen, rho, eps_thomsen, delta_thomsen, gamma_thomsen):
    """VTI stiffness from Thomsen parameters (parameters.py:71-107)."""
    C33 = alpha_thomsen ** 2 * rho
    C44 = beta_thomsen ** 2 * rho
    C11 = C33 * (1 + 2 * eps_thomsen)
    C66 = C44 * (1 + 2 * gamma_thomsen)
    C13 = torch.sqrt(2 * C33 * (C33 - C44) * delta_thomsen + (C33 - C44) ** 2) - C44
    return C11, C13, C33, C44, C66


def elastic_moduli_to_thomsen(C11, C13, C33, C44, C66, rho):
    """Inverse map (parameters.py:109-133)."""
    alpha = torch.sqrt(C33 / rho)
    beta = torch.sqrt(C44 / rho)
    eps = (C11 - C33) / (2 * C33)
    gamma = (C66 - C44) / (2 * C44)
    delta = ((C13 + C44) ** 2 - (C33 - C44) ** 2) / (2 * C33 * (C33 - C44))
    return alpha, beta, eps, delta, gamma


def _ti_fill(CC):
    C11, C12, C13, C14, C15, C16, C22, C23, C24, C25, C26, C33, C34, C35, C36, C44, C45, C46, C55, C56, C66 = CC
    return [C11, (C11 - 2 * C66).clone(), C13, C14, C15, C16, C11.clone(), C13.clone(), C24, C25, C26,
            C33, C34, C35, C36, C44, C45, C46, C44.clone(), C56, C66]


def elastic_moduli_for_isotropic(CC):
    """C22 = C11, C55 = C44, C23 = C13, C12 = C11 - 2 C66 (parameters.py:135-154)."""
    return _ti_fill(CC)


def elastic_moduli_for_TI(CC, anisotropic_type="VTI"):
    """VTI fill as for the isotropic case; HTI permutes axes (parameters.py:156-181)."""
    CC = _ti_fill(CC)
    if anisotropic_type.lower() in ["hti"]:
        C11, C12, C13, C14, C15, C16, C22, C23, C24, C25, C26, C33, C34, C35, C36, C44, C45, C46, C55, C56, C66 = CC
        CC = [C33, C13, C13, C14, C15, C16, C11, C12, C24, C25, C26, C11, C34, C35, C36, C66, C45, C46, C55,
              C56, C55]
    return CC


def _five_point(m, nx, nz):
    # the reference's 5-term average counts m[2:, 1:-1] twice (parameters.py:202-212)
    return 0.2 * (m[1:nz - 1, 1:nx - 1] + m[2:nz, 1:nx - 1] + m[1:nz - 1, 2:nx]
                  + m[2:nz, 1:nx - 1] + m[2:nz, 2:nx])


def parameter_staggered_grid(mu, b, C44, C55, C66, nx, nz):
    """bx (nz, nx-1), bz (nz-1, nx) two-point averages; shear moduli (nz-2, nx-2) (parameters.py:184-213)."""
    bx = 0.5 * (b[:, 0:nx - 1] + b[:, 1:nx])
    bz = 0.5 * (b[0:nz - 1, :] + b[1:nz, :])
    return bx, bz, _five_point(mu, nx, nz), _five_point(C44, nx, nz), _five_point(C55, nx, nz), \
        _five_point(C66, nx, nz)
