"""Oracle: the reference acoustic propagator path in eager PyTorch (test infrastructure only).

Restates ADFWI/propagator/acoustic_kernels.py:16-301 with the same update
order, index windows, fp32 association of every slice expression, chunked
reentrant checkpointing and energy-map accumulation (including the quirks at
acoustic_kernels.py:174 and :287-288).  Gradients come from torch autograd, as
in the reference; this is also the "plain autograd path" timed beside the HIP
path.
"""
from typing import Dict, Tuple

import torch
from torch.utils.checkpoint import checkpoint

C1 = 9.0 / 8.0
C2 = -1.0 / 24.0


def pad_edge(v: torch.Tensor, nabc: int, nz: int, nx: int) -> torch.Tensor:
    """Edge-replicating pad (nz, nx) -> (nz+2nabc, nx+2nabc), always float32.

    acoustic_kernels.py:16-37: the destination is ``torch.zeros`` without a
    dtype, so the result is float32 whatever the input dtype.
    """
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    out = torch.zeros((Z, X), device=v.device)
    out[nabc:Z - nabc, nabc:X - nabc] = v
    out[:nabc, nabc:nabc + nx] = out[nabc, nabc:nabc + nx].expand(nabc, -1)
    out[Z - nabc:, nabc:nabc + nx] = out[Z - nabc - 1, nabc:nabc + nx].expand(nabc, -1)
    out[:, :nabc] = out[:, [nabc]].expand(-1, nabc)
    out[:, X - nabc:] = out[:, [X - nabc - 1]].expand(-1, nabc)
    return out


@torch.jit.script
def _steps(nx: int, nz: int, dt: float, nabc: int, free_surface: bool,
           src_x: torch.Tensor, src_z: torch.Tensor, src_v: torch.Tensor,
           rcv_x: torch.Tensor, rcv_z: torch.Tensor,
           omk1: torch.Tensor, a1: torch.Tensor, omk2: torch.Tensor, a2: torch.Tensor,
           omk3: torch.Tensor, c1: float, c2: float,
           p: torch.Tensor, u: torch.Tensor, w: torch.Tensor
           ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor,
                      torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """``src_v.shape[-1]`` time steps of acoustic_kernels.py:113-174.

    omk1/omk2/omk3 are (1 - kappa) per cell; the reference forms them inside the
    loop with the same single fp32 subtraction.
    """
    p = p.clone()
    u = u.clone()
    w = w.clone()
    S = p.shape[0]
    Z = p.shape[1]
    X = p.shape[2]
    nt = src_v.shape[-1]
    R = rcv_x.shape[0]
    fs = nabc if free_surface else 1
    rec_p = torch.zeros((S, nt, R), dtype=p.dtype, device=p.device)
    rec_u = torch.zeros((S, nt, R), dtype=p.dtype, device=p.device)
    rec_w = torch.zeros((S, nt, R), dtype=p.dtype, device=p.device)
    e_p = torch.zeros((nz, nx), dtype=p.dtype, device=p.device)
    e_u = torch.zeros((nz, nx), dtype=p.dtype, device=p.device)
    e_w = torch.zeros((nz, nx), dtype=p.dtype, device=p.device)
    shots = torch.arange(S, device=p.device)
    for it in range(nt):
        # pressure on rows [fs+1, Z-3] x cols [2, X-3]
        rp0, rp1 = fs + 1, Z - 2
        p[:, rp0:rp1, 2:X - 2] = omk1[rp0:rp1, 2:X - 2] * p[:, rp0:rp1, 2:X - 2] - a1[rp0:rp1, 2:X - 2] * (
            c1 * (u[:, rp0:rp1, 2:X - 2] - u[:, rp0:rp1, 1:X - 3]
                  + w[:, rp0:rp1, 2:X - 2] - w[:, rp0 - 1:rp1 - 1, 2:X - 2])
            + c2 * (u[:, rp0:rp1, 3:X - 1] - u[:, rp0:rp1, 0:X - 4]
                    + w[:, rp0 + 1:rp1 + 1, 2:X - 2] - w[:, rp0 - 2:rp1 - 2, 2:X - 2]))
        # source term, then the free-surface mirror of pressure
        if src_v.dim() == 1:
            inj = dt * src_v[it]
        else:
            inj = dt * src_v[:, it]
        p[shots, src_z, src_x] = p[shots, src_z, src_x] + inj
        if free_surface:
            p[:, fs - 1, :] = -p[:, fs + 1, :]
        # horizontal velocity on rows [fs, Z-2] x cols [1, X-3]
        u[:, fs:Z - 1, 1:X - 2] = omk2[fs:Z - 1, 1:X - 2] * u[:, fs:Z - 1, 1:X - 2] - a2[fs:Z - 1, 1:X - 2] * (
            c1 * (p[:, fs:Z - 1, 2:X - 1] - p[:, fs:Z - 1, 1:X - 2])
            + c2 * (p[:, fs:Z - 1, 3:X] - p[:, fs:Z - 1, 0:X - 3]))
        # vertical velocity on rows [fs, Z-3] x cols [1, X-2]
        w[:, fs:Z - 2, 1:X - 1] = omk3[fs:Z - 2, 1:X - 1] * w[:, fs:Z - 2, 1:X - 1] - a2[fs:Z - 2, 1:X - 1] * (
            c1 * (p[:, fs + 1:Z - 1, 1:X - 1] - p[:, fs:Z - 2, 1:X - 1])
            + c2 * (p[:, fs + 2:Z, 1:X - 1] - p[:, fs - 1:Z - 3, 1:X - 1]))
        if free_surface:
            w[:, fs - 1, :] = w[:, fs, :]
        rec_p[:, it, :] = p[:, rcv_z, rcv_x]
        rec_u[:, it, :] = u[:, rcv_z, rcv_x]
        rec_w[:, it, :] = w[:, rcv_z, rcv_x]
        e_p = e_p + torch.sum(p * p, dim=0)[nabc:nabc + nz, nabc:nabc + nx].detach()
        e_u = e_u + torch.sum(u * u, dim=0)[nabc:nabc + nz, nabc:nabc + nx].detach()
        # the reference adds w^2 to the *u* accumulator here (acoustic_kernels.py:174)
        e_w = e_u + torch.sum(w * w, dim=0)[nabc:nabc + nz, nabc:nabc + nx].detach()
    return p, u, w, rec_p, rec_u, rec_w, e_p, e_u, e_w


def coefficients(vp: torch.Tensor, rho: torch.Tensor, damp: torch.Tensor, nz: int, nx: int,
                 nabc: int, dt: float, dz: float, free_surface: bool):
    """Padded coefficient fields (acoustic_kernels.py:225-265).

    Returns (omk1, a1, omk2, a2, omk3) with omk = 1 - kappa.
    """
    c = pad_edge(vp, nabc, nz, nx)
    den = pad_edge(rho, nabc, nz, nx)
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    fs = nabc if free_surface else 1
    # dz as a 0-d tensor on the data's device: on a GPU torch turns division by a Python number
    # into multiplication by its reciprocal; the reference divides (it runs on the CPU)
    a1 = den * c * c * dt / torch.full([], dz, dtype=den.dtype, device=den.device)
    k1 = damp * dt
    a2 = dt / (den * dz)
    k2 = torch.zeros_like(damp)
    k2[:, 1:X - 2] = 0.5 * (damp[:, 1:X - 2] + damp[:, 2:X - 1]) * dt
    k3 = torch.zeros_like(damp)
    k3[fs:Z - 2, :] = 0.5 * (damp[fs:Z - 2, :] + damp[fs + 1:Z - 1, :]) * dt
    return 1.0 - k1, a1, 1.0 - k2, a2, 1.0 - k3


def forward_kernel(nx: int, nz: int, dx: float, dz: float, nt: int, dt: float,
                   nabc: int, free_surface: bool,
                   src_x: torch.Tensor, src_z: torch.Tensor, src_n: int, src_v: torch.Tensor,
                   rcv_x: torch.Tensor, rcv_z: torch.Tensor, rcv_n: int,
                   damp: torch.Tensor, v: torch.Tensor, rho: torch.Tensor,
                   checkpoint_segments: int = 1,
                   device=torch.device('cpu'), dtype=torch.float32) -> Dict[str, torch.Tensor]:
    """Acoustic forward modelling, acoustic_kernels.py:179-301 semantics."""
    omk1, a1, omk2, a2, omk3 = coefficients(v, rho, damp, nz, nx, nabc, dt, dz, free_surface)
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    sx = src_x + nabc
    sz = src_z + nabc
    rx = rcv_x + nabc
    rz = rcv_z + nabc
    p = torch.zeros((src_n, Z, X), dtype=dtype, device=device)
    u = torch.zeros((src_n, Z, X - 1), dtype=dtype, device=device)
    w = torch.zeros((src_n, Z - 1, X), dtype=dtype, device=device)
    rec_p = torch.zeros((src_n, nt, rcv_n), dtype=dtype, device=device)
    rec_u = torch.zeros((src_n, nt, rcv_n), dtype=dtype, device=device)
    rec_w = torch.zeros((src_n, nt, rcv_n), dtype=dtype, device=device)
    f_p = torch.zeros((nz, nx), dtype=dtype, device=device)
    f_u = torch.zeros((nz, nx), dtype=dtype, device=device)
    f_w = torch.zeros((nz, nx), dtype=dtype, device=device)
    t0 = 0
    for chunk in torch.chunk(src_v, checkpoint_segments, dim=-1):
        p, u, w, cp, cu, cw, ep, eu, ew = checkpoint(
            _steps, nx, nz, dt, nabc, free_surface, sx, sz, chunk, rx, rz,
            omk1, a1, omk2, a2, omk3, C1, C2, p, u, w, use_reentrant=True)
        n = chunk.shape[-1]
        rec_p[:, t0:t0 + n] = cp
        rec_u[:, t0:t0 + n] = cu
        rec_w[:, t0:t0 + n] = cw
        # cross-chunk accumulation as at acoustic_kernels.py:286-288
        f_p = f_p + ep.detach()
        f_u = f_p + eu.detach()
        f_w = f_p + ew.detach()
        t0 += n
    return {"p": rec_p, "u": rec_u, "w": rec_w,
            "forward_wavefield_p": f_p, "forward_wavefield_u": f_u, "forward_wavefield_w": f_w}


def l2_misfit(obs: torch.Tensor, syn: torch.Tensor, dt: float = 1.0) -> torch.Tensor:
    """Waveform L2 misfit of ADFWI/fwi/misfit/L2.py:26-29: sum_{s,r} sqrt(sum_t rsd^2 dt)."""
    rsd = obs - syn
    return torch.sum(torch.sqrt(torch.sum(rsd * rsd * dt, dim=1)))
