"""Oracle: the reference elastic propagator path in eager PyTorch.

Test infrastructure only.  Restates ADFWI/propagator/elastic_kernels.py:
DiffCoef (:20-58), the 4th- and 6th-order D operators (:66-108), pad_torchSingle
(:175-216), step_forward_PML_4order (:222-423), step_forward_PML_6order (:426-584),
step_forward_ABL_4order (:585-777), step_forward_ABL_6order (:781-912) and
forward_kernel (:917-1037), keeping the
fp32 association of every expression, the split-field PML, the moment-tensor
source (-MT/2 into both splits), the free-surface rows and the per-step
overwrite of the energy maps.  For fd_order 6, 8 and 10 the reference runs
step_forward_PML_6order (step_forward_ABL_6order for the damping scheme), whose
three-term operators use the first three
coefficients of DiffCoef(fd_order // 2) on the window [fd_order // 2, Z - fd_order // 2 - 1]
with the free-surface rows at h = fd_order // 2 + 1.
"""
from typing import Dict, List, Tuple

import torch
from torch.utils.checkpoint import checkpoint


def diff_coef(order: int) -> torch.Tensor:
    """Staggered ('s' scheme) FD coefficients solved in float32 (elastic_kernels.py:20-58)."""
    a = torch.empty((order, order), dtype=torch.float32)
    for i in range(order):
        for j in range(order):
            a[i, j] = (2 * (j + 1) - 1) ** (2 * (i + 1) - 1)
    b = torch.cat((torch.tensor([1.0]), torch.zeros(order - 1)))
    return torch.linalg.inv(a) @ b


def pad_elastic(data: torch.Tensor, pml: int, fs_offset: int, free_surface: bool) -> torch.Tensor:
    """Edge-replicating pad of an array by its own shape, always float32 (elastic_kernels.py:175-216)."""
    nz, nx = data.shape
    X = nx + 2 * pml
    Z = nz + (pml + fs_offset if free_surface else 2 * pml + fs_offset)
    out = torch.zeros((Z, X), device=data.device)
    top = fs_offset if free_surface else fs_offset + pml
    out[top:top + nz, pml:pml + nx] = data
    out[:top, pml:pml + nx] = data[0, :].unsqueeze(0).expand(top, -1)
    out[Z - pml:, pml:pml + nx] = data[-1, :].unsqueeze(0).expand(pml, -1)
    out[:, :pml] = out[:, pml:pml + 1].expand(-1, pml)
    out[:, X - pml:] = out[:, X - pml - 1:X - pml].expand(-1, pml)
    return out


@torch.jit.script
def _dxf(a: torch.Tensor, c: torch.Tensor, Z: int, X: int) -> torch.Tensor:
    return c[0] * (a[:, 2:Z - 2, 3:X - 1] - a[:, 2:Z - 2, 2:X - 2]) + \
        c[1] * (a[:, 2:Z - 2, 4:X] - a[:, 2:Z - 2, 1:X - 3])


@torch.jit.script
def _dxb(a: torch.Tensor, c: torch.Tensor, Z: int, X: int) -> torch.Tensor:
    return c[0] * (a[:, 2:Z - 2, 2:X - 2] - a[:, 2:Z - 2, 1:X - 3]) + \
        c[1] * (a[:, 2:Z - 2, 3:X - 1] - a[:, 2:Z - 2, 0:X - 4])


@torch.jit.script
def _dzf(a: torch.Tensor, c: torch.Tensor, Z: int, X: int) -> torch.Tensor:
    return c[0] * (a[:, 3:Z - 1, 2:X - 2] - a[:, 2:Z - 2, 2:X - 2]) + \
        c[1] * (a[:, 4:Z, 2:X - 2] - a[:, 1:Z - 3, 2:X - 2])


@torch.jit.script
def _dzb(a: torch.Tensor, c: torch.Tensor, Z: int, X: int) -> torch.Tensor:
    return c[0] * (a[:, 2:Z - 2, 2:X - 2] - a[:, 1:Z - 3, 2:X - 2]) + \
        c[1] * (a[:, 3:Z - 1, 2:X - 2] - a[:, 0:Z - 4, 2:X - 2])


@torch.jit.script
def _steps(free_surface: bool, nx: int, nz: int, dx: float, dz: float, nabc: int,
           src_x: torch.Tensor, src_z: torch.Tensor, dt: float, src_v: torch.Tensor, mt: torch.Tensor,
           rcv_x: torch.Tensor, rcv_z: torch.Tensor, bcx: torch.Tensor, bcz: torch.Tensor,
           c11: torch.Tensor, c13: torch.Tensor, c15: torch.Tensor, c33: torch.Tensor,
           c35: torch.Tensor, c55: torch.Tensor, bx: torch.Tensor, bz: torch.Tensor, fdc: torch.Tensor,
           txx_x: torch.Tensor, txx_z: torch.Tensor, tzz_x: torch.Tensor, tzz_z: torch.Tensor,
           txz_x: torch.Tensor, txz_z: torch.Tensor, txx: torch.Tensor, tzz: torch.Tensor, txz: torch.Tensor,
           vx_x: torch.Tensor, vx_z: torch.Tensor, vz_x: torch.Tensor, vz_z: torch.Tensor,
           vx: torch.Tensor, vz: torch.Tensor) -> List[torch.Tensor]:
    """``src_v.shape[-1]`` steps of elastic_kernels.py:339-418 (fd order 4, PML)."""
    S = vx.shape[0]
    Z = vx.shape[1]
    X = vx.shape[2]
    nt = src_v.shape[-1]
    R = rcv_x.shape[0]
    txx_x, txx_z, tzz_x, tzz_z = txx_x.clone(), txx_z.clone(), tzz_x.clone(), tzz_z.clone()
    txz_x, txz_z, txx, tzz, txz = txz_x.clone(), txz_z.clone(), txx.clone(), tzz.clone(), txz.clone()
    vx_x, vx_z, vz_x, vz_z = vx_x.clone(), vx_z.clone(), vz_x.clone(), vz_z.clone()
    vx, vz = vx.clone(), vz.clone()
    recs = [torch.zeros((S, nt, R), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    ens = [torch.zeros((nz, nx), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    # window rows [2, Z-3], cols [2, X-3]
    W0, WZ, WX = 2, Z - 2, X - 2
    bxw = bcx[W0:WZ, W0:WX]
    bzw = bcz[W0:WZ, W0:WX]
    pxd = 1 + 0.5 * dt * bxw
    pxn = 1 - 0.5 * dt * bxw
    pzd = 1 + 0.5 * dt * bzw
    pzn = 1 - 0.5 * dt * bzw
    dt_dx = dt / dx
    dt_dz = dt / dz
    # divisors as 0-d tensors on the data's device: on a GPU torch turns division by a Python
    # number into multiplication by its reciprocal; the reference divides (it runs on the CPU)
    dxs = torch.full([], dx, dtype=vx.dtype, device=vx.device)
    dzs = torch.full([], dz, dtype=vx.dtype, device=vx.device)
    pxi = 1.0 / pxd
    pzi = 1.0 / pzd
    shots = torch.arange(S, device=vx.device)
    half = mt / 2
    row0 = 2 if free_surface else 2 + nabc
    C11 = c11[W0:WZ, W0:WX]
    C13 = c13[W0:WZ, W0:WX]
    C15 = c15[W0:WZ, W0:WX]
    C33 = c33[W0:WZ, W0:WX]
    C35 = c35[W0:WZ, W0:WX]
    C55 = c55[W0:WZ, W0:WX]
    BX = bx[W0:WZ, W0:WX]
    BZ = bz[W0:WZ, W0:WX]
    for t in range(nt):
        dxb_vx = _dxb(vx, fdc, Z, X)
        dxb_vz = _dxb(vz, fdc, Z, X)
        dzb_vx = _dzb(vx, fdc, Z, X)
        dzb_vz = _dzb(vz, fdc, Z, X)
        dxf_vx = _dxf(vx, fdc, Z, X)
        dzf_vx = _dzf(vx, fdc, Z, X)
        dxf_vz = _dxf(vz, fdc, Z, X)
        dzf_vz = _dzf(vz, fdc, Z, X)
        txx_x[:, W0:WZ, W0:WX] = (pxn * txx_x[:, W0:WZ, W0:WX] + dt_dx * (C11 * dxb_vx + C15 * dxb_vz)) * pxi
        txx_z[:, W0:WZ, W0:WX] = (pzn * txx_z[:, W0:WZ, W0:WX] + dt_dz * (C15 * dzb_vx + C13 * dzb_vz)) * pzi
        tzz_x[:, W0:WZ, W0:WX] = (pxn * tzz_x[:, W0:WZ, W0:WX] + dt_dx * (C13 * dxb_vx + C35 * dxb_vz)) * pxi
        tzz_z[:, W0:WZ, W0:WX] = (pzn * tzz_z[:, W0:WZ, W0:WX] + dt_dz * (C35 * dzb_vx + C33 * dzb_vz)) * pzi
        txz_x[:, W0:WZ, W0:WX] = (pxn * txz_x[:, W0:WZ, W0:WX] + dt_dx * (C15 * dxf_vx + C55 * dxf_vz)) * pxi
        txz_z[:, W0:WZ, W0:WX] = (pzn * txz_z[:, W0:WZ, W0:WX] + dt_dz * (C55 * dzf_vx + C35 * dzf_vz)) * pzi
        # moment-tensor source: -MT/2 into each split, no dt factor
        sv = src_v[shots, t]
        txx_x[shots, src_z, src_x] += -half[:, 0, 0] * sv
        txx_z[shots, src_z, src_x] += -half[:, 0, 0] * sv
        tzz_x[shots, src_z, src_x] += -half[:, 2, 2] * sv
        tzz_z[shots, src_z, src_x] += -half[:, 2, 2] * sv
        txz_x[shots, src_z, src_x] += -half[:, 0, 2] * sv
        txz_z[shots, src_z, src_x] += -half[:, 0, 2] * sv
        txx[:] = txx_x + txx_z
        tzz[:] = tzz_x + tzz_z
        txz[:] = txz_x + txz_z
        if free_surface:
            tzz[:, 2, :] = 0
            tzz[:, 1, :] = -tzz[:, 3, :]
            txz[:, 1, :] = -txz[:, 2, :]
            txz[:, 0, :] = -txz[:, 3, :]
        dxf_txx = _dxf(txx, fdc, Z, X)
        dzb_txz = _dzb(txz, fdc, Z, X)
        dxb_txz = _dxb(txz, fdc, Z, X)
        dzf_tzz = _dzf(tzz, fdc, Z, X)
        vx_x[:, W0:WZ, W0:WX] = (pxn * vx_x[:, W0:WZ, W0:WX] + dt * BX * dxf_txx / dxs) / pxd
        vx_z[:, W0:WZ, W0:WX] = (pzn * vx_z[:, W0:WZ, W0:WX] + dt * BX * dzb_txz / dzs) / pzd
        vz_x[:, W0:WZ, W0:WX] = (pxn * vz_x[:, W0:WZ, W0:WX] + dt * BZ * dxb_txz / dxs) / pxd
        vz_z[:, W0:WZ, W0:WX] = (pzn * vz_z[:, W0:WZ, W0:WX] + dt * BZ * dzf_tzz / dzs) / pzd
        vx[:] = vx_x + vx_z
        vz[:] = vz_x + vz_z
        if free_surface:
            vz[:, 1, W0:WX] = vz[:, 2, W0:WX]
            vx[:, 1, W0:WX] = vz[:, 1, W0 + 1:WX + 1] - vz[:, 1, W0:WX] + vz[:, 2, W0 + 1:WX + 1] \
                - vz[:, 2, W0:WX] + vx[:, 3, W0:WX]
            vz[:, 0, W0:WX] = vz[:, 1, W0:WX]
        recs[0][:, t, :] = txx[:, rcv_z, rcv_x]
        recs[1][:, t, :] = tzz[:, rcv_z, rcv_x]
        recs[2][:, t, :] = txz[:, rcv_z, rcv_x]
        recs[3][:, t, :] = vx[:, rcv_z, rcv_x]
        recs[4][:, t, :] = vz[:, rcv_z, rcv_x]
        ens[0] = torch.sum(txx * txx, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[1] = torch.sum(tzz * tzz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[2] = torch.sum(txz * txz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[3] = torch.sum(vx * vx, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[4] = torch.sum(vz * vz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
    return [txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z, txx, tzz, txz,
            vx_x, vx_z, vz_x, vz_z, vx, vz] + recs + ens


@torch.jit.script
def _steps_abl(free_surface: bool, nx: int, nz: int, dx: float, dz: float, nabc: int,
               src_x: torch.Tensor, src_z: torch.Tensor, dt: float, src_v: torch.Tensor, mt: torch.Tensor,
               rcv_x: torch.Tensor, rcv_z: torch.Tensor, damp: torch.Tensor,
               c11: torch.Tensor, c13: torch.Tensor, c15: torch.Tensor, c33: torch.Tensor,
               c35: torch.Tensor, c55: torch.Tensor, bx: torch.Tensor, bz: torch.Tensor, fdc: torch.Tensor,
               txx: torch.Tensor, tzz: torch.Tensor, txz: torch.Tensor,
               vx: torch.Tensor, vz: torch.Tensor) -> List[torch.Tensor]:
    """``src_v.shape[-1]`` steps of step_forward_ABL_4order (elastic_kernels.py:585-777): unsplit
    stress and velocity updates, -MT/3 source (:703, :727-735), stress and velocity free-surface
    rows kept in the state (:738-742, :753-756), multiplicative damping of vx, vz (:759-760)."""
    S = vx.shape[0]
    Z = vx.shape[1]
    X = vx.shape[2]
    nt = src_v.shape[-1]
    R = rcv_x.shape[0]
    txx, tzz, txz, vx, vz = txx.clone(), tzz.clone(), txz.clone(), vx.clone(), vz.clone()
    recs = [torch.zeros((S, nt, R), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    ens = [torch.zeros((nz, nx), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    W0, WZ, WX = 2, Z - 2, X - 2
    shots = torch.arange(S, device=vx.device)
    scaling = -1.0 / 3
    # divisors as 0-d tensors on the data's device: on a GPU torch turns division by a Python
    # number into multiplication by its reciprocal; the reference divides (it runs on the CPU)
    dxs = torch.full([], dx, dtype=vx.dtype, device=vx.device)
    dzs = torch.full([], dz, dtype=vx.dtype, device=vx.device)
    row0 = 2 if free_surface else 2 + nabc
    C11 = c11[W0:WZ, W0:WX]
    C13 = c13[W0:WZ, W0:WX]
    C15 = c15[W0:WZ, W0:WX]
    C33 = c33[W0:WZ, W0:WX]
    C35 = c35[W0:WZ, W0:WX]
    C55 = c55[W0:WZ, W0:WX]
    BX = bx[W0:WZ, W0:WX]
    BZ = bz[W0:WZ, W0:WX]
    for t in range(nt):
        dxb_vx = _dxb(vx, fdc, Z, X)
        dxb_vz = _dxb(vz, fdc, Z, X)
        dzb_vx = _dzb(vx, fdc, Z, X)
        dzb_vz = _dzb(vz, fdc, Z, X)
        dxf_vx = _dxf(vx, fdc, Z, X)
        dzf_vx = _dzf(vx, fdc, Z, X)
        dxf_vz = _dxf(vz, fdc, Z, X)
        dzf_vz = _dzf(vz, fdc, Z, X)
        txx[:, W0:WZ, W0:WX] = txx[:, W0:WZ, W0:WX] + dt * ((C11 * dxb_vx + C15 * dxb_vz) / dxs
                                                            + (C15 * dzb_vx + C13 * dzb_vz) / dzs)
        tzz[:, W0:WZ, W0:WX] = tzz[:, W0:WZ, W0:WX] + dt * ((C13 * dxb_vx + C35 * dxb_vz) / dxs
                                                            + (C35 * dzb_vx + C33 * dzb_vz) / dzs)
        txz[:, W0:WZ, W0:WX] = txz[:, W0:WZ, W0:WX] + dt * ((C15 * dxf_vx + C55 * dxf_vz) / dxs
                                                            + (C55 * dzf_vx + C35 * dzf_vz) / dzs)
        sv = src_v[shots, t]
        txx[shots, src_z, src_x] += scaling * mt[:, 0, 0] * sv
        tzz[shots, src_z, src_x] += scaling * mt[:, 2, 2] * sv
        txz[shots, src_z, src_x] += scaling * mt[:, 0, 2] * sv
        if free_surface:
            tzz[:, 2, :] = 0
            tzz[:, 1, :] = -tzz[:, 3, :]
            txz[:, 1, :] = -txz[:, 2, :]
            txz[:, 0, :] = -txz[:, 3, :]
        dxf_txx = _dxf(txx, fdc, Z, X)
        dzb_txz = _dzb(txz, fdc, Z, X)
        dxb_txz = _dxb(txz, fdc, Z, X)
        dzf_tzz = _dzf(tzz, fdc, Z, X)
        vx[:, W0:WZ, W0:WX] += dt * BX * (dxf_txx / dxs + dzb_txz / dzs)
        vz[:, W0:WZ, W0:WX] += dt * BZ * (dxb_txz / dxs + dzf_tzz / dzs)
        if free_surface:
            vz[:, 1, W0:WX] = vz[:, 2, W0:WX]
            vx[:, 1, W0:WX] = vz[:, 1, W0 + 1:WX + 1] - vz[:, 1, W0:WX] + vz[:, 2, W0 + 1:WX + 1] \
                - vz[:, 2, W0:WX] + vx[:, 3, W0:WX]
            vz[:, 0, W0:WX] = vz[:, 1, W0:WX]
        vx = vx * damp
        vz = vz * damp
        recs[0][:, t, :] = txx[:, rcv_z, rcv_x]
        recs[1][:, t, :] = tzz[:, rcv_z, rcv_x]
        recs[2][:, t, :] = txz[:, rcv_z, rcv_x]
        recs[3][:, t, :] = vx[:, rcv_z, rcv_x]
        recs[4][:, t, :] = vz[:, rcv_z, rcv_x]
        ens[0] = torch.sum(txx * txx, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[1] = torch.sum(tzz * tzz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[2] = torch.sum(txz * txz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[3] = torch.sum(vx * vx, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
        ens[4] = torch.sum(vz * vz, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
    return [txx, tzz, txz, vx, vz] + recs + ens


def _steps_abl_tuple(*args):
    return tuple(_steps_abl(*args))


def _steps_tuple(*args):
    """Tuple-returning wrapper: checkpoint() tracks tensors returned in a tuple, not a list."""
    return tuple(_steps(*args))


def _steps6(free_surface: bool, nx: int, nz: int, dx: float, dz: float, nabc: int, nn: int,
            src_x, src_z, dt: float, src_v, mt, rcv_x, rcv_z, bcx, bcz, c11, c13, c15, c33, c35, c55, bx, bz,
            fdc, txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z, vx, vz):
    """``src_v.shape[-1]`` steps of step_forward_PML_6order (elastic_kernels.py:426-584): the
    order-4 step with three-term operators (:85-108), the window [nn, Z-nn-1] and the
    free-surface rows at h = nn + 1."""
    S, Z, X = vx.shape
    nt = src_v.shape[-1]
    R = rcv_x.shape[0]
    txx_x, txx_z, tzz_x, tzz_z = txx_x.clone(), txx_z.clone(), tzz_x.clone(), tzz_z.clone()
    txz_x, txz_z, txx, tzz, txz = txz_x.clone(), txz_z.clone(), txx.clone(), tzz.clone(), txz.clone()
    vx_x, vx_z, vz_x, vz_z, vx, vz = vx_x.clone(), vx_z.clone(), vz_x.clone(), vz_z.clone(), vx.clone(), vz.clone()
    recs = [torch.zeros((S, nt, R), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    ens = [torch.zeros((nz, nx), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    I, J = slice(nn, Z - nn), slice(nn, X - nn)
    c0, c1, c2 = fdc[0], fdc[1], fdc[2]

    def sh(a, dz_, dx_):   # a[:, i + dz_, j + dx_] over the window
        return a[:, nn + dz_:Z - nn + dz_, nn + dx_:X - nn + dx_]

    def dxf(a):
        return c0 * (sh(a, 0, 1) - sh(a, 0, 0)) + c1 * (sh(a, 0, 2) - sh(a, 0, -1)) + c2 * (sh(a, 0, 3) - sh(a, 0, -2))

    def dzf(a):
        return c0 * (sh(a, 1, 0) - sh(a, 0, 0)) + c1 * (sh(a, 2, 0) - sh(a, -1, 0)) + c2 * (sh(a, 3, 0) - sh(a, -2, 0))

    def dxb(a):
        return c0 * (sh(a, 0, 0) - sh(a, 0, -1)) + c1 * (sh(a, 0, 1) - sh(a, 0, -2)) + c2 * (sh(a, 0, 2) - sh(a, 0, -3))

    def dzb(a):
        return c0 * (sh(a, 0, 0) - sh(a, -1, 0)) + c1 * (sh(a, 1, 0) - sh(a, -2, 0)) + c2 * (sh(a, 2, 0) - sh(a, -3, 0))
    bxw, bzw = bcx[I, J], bcz[I, J]
    pxd, pxn = 1 + 0.5 * dt * bxw, 1 - 0.5 * dt * bxw
    pzd, pzn = 1 + 0.5 * dt * bzw, 1 - 0.5 * dt * bzw
    dt_dx, dt_dz = dt / dx, dt / dz
    # divisors as 0-d tensors on the data's device: on a GPU torch turns division by a Python
    # number into multiplication by its reciprocal; the reference divides (it runs on the CPU)
    dxs = torch.full([], dx, dtype=vx.dtype, device=vx.device)
    dzs = torch.full([], dz, dtype=vx.dtype, device=vx.device)
    pxi, pzi = 1.0 / pxd, 1.0 / pzd
    shots = torch.arange(S, device=vx.device)
    half = mt / 2
    h = nn + 1
    row0 = nn if free_surface else nn + nabc
    C11, C13, C15, C33, C35, C55 = (a[I, J] for a in (c11, c13, c15, c33, c35, c55))
    BX, BZ = bx[I, J], bz[I, J]
    for t in range(nt):
        dxb_vx, dxb_vz, dzb_vx, dzb_vz = dxb(vx), dxb(vz), dzb(vx), dzb(vz)
        dxf_vx, dzf_vx, dxf_vz, dzf_vz = dxf(vx), dzf(vx), dxf(vz), dzf(vz)
        txx_x[:, I, J] = (pxn * txx_x[:, I, J] + dt_dx * (C11 * dxb_vx + C15 * dxb_vz)) * pxi
        txx_z[:, I, J] = (pzn * txx_z[:, I, J] + dt_dz * (C15 * dzb_vx + C13 * dzb_vz)) * pzi
        tzz_x[:, I, J] = (pxn * tzz_x[:, I, J] + dt_dx * (C13 * dxb_vx + C35 * dxb_vz)) * pxi
        tzz_z[:, I, J] = (pzn * tzz_z[:, I, J] + dt_dz * (C35 * dzb_vx + C33 * dzb_vz)) * pzi
        txz_x[:, I, J] = (pxn * txz_x[:, I, J] + dt_dx * (C15 * dxf_vx + C55 * dxf_vz)) * pxi
        txz_z[:, I, J] = (pzn * txz_z[:, I, J] + dt_dz * (C55 * dzf_vx + C35 * dzf_vz)) * pzi
        sv = src_v[shots, t]
        txx_x[shots, src_z, src_x] += -half[:, 0, 0] * sv
        txx_z[shots, src_z, src_x] += -half[:, 0, 0] * sv
        tzz_x[shots, src_z, src_x] += -half[:, 2, 2] * sv
        tzz_z[shots, src_z, src_x] += -half[:, 2, 2] * sv
        txz_x[shots, src_z, src_x] += -half[:, 0, 2] * sv
        txz_z[shots, src_z, src_x] += -half[:, 0, 2] * sv
        txx[:] = txx_x + txx_z
        tzz[:] = tzz_x + tzz_z
        txz[:] = txz_x + txz_z
        if free_surface:
            tzz[:, h - 1, :] = 0
            tzz[:, h - 2, :] = -tzz[:, h, :]
            txz[:, h - 2, :] = -txz[:, h - 1, :]
            txz[:, h - 3, :] = -txz[:, h, :]
        dxf_txx, dzb_txz, dxb_txz, dzf_tzz = dxf(txx), dzb(txz), dxb(txz), dzf(tzz)
        vx_x[:, I, J] = (pxn * vx_x[:, I, J] + dt * BX * dxf_txx / dxs) / pxd
        vx_z[:, I, J] = (pzn * vx_z[:, I, J] + dt * BX * dzb_txz / dzs) / pzd
        vz_x[:, I, J] = (pxn * vz_x[:, I, J] + dt * BZ * dxb_txz / dxs) / pxd
        vz_z[:, I, J] = (pzn * vz_z[:, I, J] + dt * BZ * dzf_tzz / dzs) / pzd
        vx[:] = vx_x + vx_z
        vz[:] = vz_x + vz_z
        if free_surface:
            J1 = slice(nn + 1, X - nn + 1)
            vz[:, h - 2, J] = vz[:, h - 1, J]
            vx[:, h - 2, J] = vz[:, h - 2, J1] - vz[:, h - 2, J] + vz[:, h - 1, J1] - vz[:, h - 1, J] + vx[:, h, J]
            vz[:, h - 3, J] = vz[:, h - 2, J]
        for k, f in enumerate((txx, tzz, txz, vx, vz)):
            recs[k][:, t, :] = f[:, rcv_z, rcv_x]
            ens[k] = torch.sum(f * f, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
    return (txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z, vx, vz,
            *recs, *ens)


def _steps_abl6(free_surface: bool, nx: int, nz: int, dx: float, dz: float, nabc: int, nn: int,
                src_x, src_z, dt: float, src_v, mt, rcv_x, rcv_z, damp, c11, c13, c15, c33, c35, c55, bx, bz,
                fdc, txx, tzz, txz, vx, vz):
    """``src_v.shape[-1]`` steps of step_forward_ABL_6order (elastic_kernels.py:781-912): the
    damping step of :585-777 with three-term operators, the window [nn, Z-nn-1] and the
    free-surface rows at h = nn + 1, kept in the state."""
    S, Z, X = vx.shape
    nt = src_v.shape[-1]
    R = rcv_x.shape[0]
    txx, tzz, txz, vx, vz = txx.clone(), tzz.clone(), txz.clone(), vx.clone(), vz.clone()
    recs = [torch.zeros((S, nt, R), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    ens = [torch.zeros((nz, nx), dtype=vx.dtype, device=vx.device) for _ in range(5)]
    I, J = slice(nn, Z - nn), slice(nn, X - nn)
    c0, c1, c2 = fdc[0], fdc[1], fdc[2]

    def sh(a, dz_, dx_):
        return a[:, nn + dz_:Z - nn + dz_, nn + dx_:X - nn + dx_]

    def dxf(a):
        return c0 * (sh(a, 0, 1) - sh(a, 0, 0)) + c1 * (sh(a, 0, 2) - sh(a, 0, -1)) + c2 * (sh(a, 0, 3) - sh(a, 0, -2))

    def dzf(a):
        return c0 * (sh(a, 1, 0) - sh(a, 0, 0)) + c1 * (sh(a, 2, 0) - sh(a, -1, 0)) + c2 * (sh(a, 3, 0) - sh(a, -2, 0))

    def dxb(a):
        return c0 * (sh(a, 0, 0) - sh(a, 0, -1)) + c1 * (sh(a, 0, 1) - sh(a, 0, -2)) + c2 * (sh(a, 0, 2) - sh(a, 0, -3))

    def dzb(a):
        return c0 * (sh(a, 0, 0) - sh(a, -1, 0)) + c1 * (sh(a, 1, 0) - sh(a, -2, 0)) + c2 * (sh(a, 2, 0) - sh(a, -3, 0))
    shots = torch.arange(S, device=vx.device)
    scaling = -1.0 / 3
    # divisors as 0-d tensors on the data's device: on a GPU torch turns division by a Python
    # number into multiplication by its reciprocal; the reference divides (it runs on the CPU)
    dxs = torch.full([], dx, dtype=vx.dtype, device=vx.device)
    dzs = torch.full([], dz, dtype=vx.dtype, device=vx.device)
    h = nn + 1
    row0 = nn if free_surface else nn + nabc
    C11, C13, C15, C33, C35, C55 = (a[I, J] for a in (c11, c13, c15, c33, c35, c55))
    BX, BZ = bx[I, J], bz[I, J]
    for t in range(nt):
        dxb_vx, dxb_vz, dzb_vx, dzb_vz = dxb(vx), dxb(vz), dzb(vx), dzb(vz)
        dxf_vx, dzf_vx, dxf_vz, dzf_vz = dxf(vx), dzf(vx), dxf(vz), dzf(vz)
        txx[:, I, J] = txx[:, I, J] + dt * ((C11 * dxb_vx + C15 * dxb_vz) / dxs + (C15 * dzb_vx + C13 * dzb_vz) / dzs)
        tzz[:, I, J] = tzz[:, I, J] + dt * ((C13 * dxb_vx + C35 * dxb_vz) / dxs + (C35 * dzb_vx + C33 * dzb_vz) / dzs)
        txz[:, I, J] = txz[:, I, J] + dt * ((C15 * dxf_vx + C55 * dxf_vz) / dxs + (C55 * dzf_vx + C35 * dzf_vz) / dzs)
        sv = src_v[shots, t]
        txx[shots, src_z, src_x] += scaling * mt[:, 0, 0] * sv
        tzz[shots, src_z, src_x] += scaling * mt[:, 2, 2] * sv
        txz[shots, src_z, src_x] += scaling * mt[:, 0, 2] * sv
        if free_surface:
            tzz[:, h - 1, :] = 0
            tzz[:, h - 2, :] = -tzz[:, h, :]
            txz[:, h - 2, :] = -txz[:, h - 1, :]
            txz[:, h - 3, :] = -txz[:, h, :]
        dxf_txx, dzb_txz, dxb_txz, dzf_tzz = dxf(txx), dzb(txz), dxb(txz), dzf(tzz)
        vx[:, I, J] += dt * BX * (dxf_txx / dxs + dzb_txz / dzs)
        vz[:, I, J] += dt * BZ * (dxb_txz / dxs + dzf_tzz / dzs)
        if free_surface:
            J1 = slice(nn + 1, X - nn + 1)
            vz[:, h - 2, J] = vz[:, h - 1, J]
            vx[:, h - 2, J] = vz[:, h - 2, J1] - vz[:, h - 2, J] + vz[:, h - 1, J1] - vz[:, h - 1, J] + vx[:, h, J]
            vz[:, h - 3, J] = vz[:, h - 2, J]
        vx = vx * damp
        vz = vz * damp
        for k, f in enumerate((txx, tzz, txz, vx, vz)):
            recs[k][:, t, :] = f[:, rcv_z, rcv_x]
            ens[k] = torch.sum(f * f, dim=0)[row0:row0 + nz, nabc:nabc + nx].detach()
    return (txx, tzz, txz, vx, vz, *recs, *ens)


def padded_fields(bx, bz, CC, bcx, bcz, nabc: int, free_surface: bool, fd_order: int = 4):
    """Pads the coefficient fields the PML kernel reads (elastic_kernels.py:935-953)."""
    fso = fd_order // 2
    c11, c13, c15, c33, c35, c55 = CC[0], CC[2], CC[4], CC[11], CC[13], CC[18]
    out = {k: pad_elastic(a, nabc, fso, free_surface) for k, a in
           (("c11", c11), ("c13", c13), ("c15", c15), ("c33", c33), ("c35", c35), ("c55", c55),
            ("bx", bx), ("bz", bz))}
    out["bcx"] = None if bcx is None else pad_elastic(bcx, 0, fso, free_surface)
    out["bcz"] = None if bcz is None else pad_elastic(bcz, 0, fso, free_surface)
    return out


def forward_kernel(nx: int, nz: int, dx: float, dz: float, nt: int, dt: float,
                   nabc: int, free_surface: bool,
                   src_x: torch.Tensor, src_z: torch.Tensor, src_n: int, src_v: torch.Tensor, MT: torch.Tensor,
                   rcv_x: torch.Tensor, rcv_z: torch.Tensor, rcv_n: int,
                   abc_type: str, bcx: torch.Tensor, bcz: torch.Tensor, damp,
                   lamu, lam, bx: torch.Tensor, bz: torch.Tensor, CC: List[torch.Tensor],
                   fd_order: int = 4, n_segments: int = 1,
                   device=torch.device("cpu"), dtype=torch.float32) -> Dict[str, torch.Tensor]:
    """Elastic forward modelling, elastic_kernels.py:917-1037 semantics: the split-field PML
    branch for abc_type 'pml' (any case, :948, :976), the damping (ABL) branch otherwise; order 4,
    or the three-term steps (:426-584, :781-912) for orders 6, 8, 10."""
    pml = abc_type.lower() == "pml"
    if fd_order not in (4, 6, 8, 10):
        raise NotImplementedError("oracle covers fd_order 4, 6, 8 and 10")
    fso = fd_order // 2
    X = nx + 2 * nabc
    Z = nz + (nabc + fso if free_surface else 2 * nabc + fso)
    f = padded_fields(bx, bz, CC, bcx if pml else None, bcz if pml else None, nabc, free_surface, fd_order)
    dmp = None if pml else pad_elastic(damp, 0, fso, free_surface)
    fdc = diff_coef(fso)
    sx = src_x + nabc
    sz = src_z + fso if free_surface else src_z + nabc + fso
    rx = rcv_x + nabc
    rz = rcv_z + fso if free_surface else rcv_z + nabc + fso
    nstate = 15 if pml else 5
    state = [torch.zeros((src_n, Z, X), dtype=dtype, device=device) for _ in range(nstate)]
    recs = [torch.zeros((src_n, nt, rcv_n), dtype=dtype, device=device) for _ in range(5)]
    ens = [torch.zeros((nz, nx), dtype=dtype, device=device) for _ in range(5)]
    t0 = 0
    for chunk in torch.chunk(src_v, n_segments, dim=-1):
        if pml and fd_order != 4:
            out = checkpoint(_steps6, free_surface, nx, nz, dx, dz, nabc, fso, sx, sz, dt, chunk, MT, rx, rz,
                             f["bcx"], f["bcz"], f["c11"], f["c13"], f["c15"], f["c33"], f["c35"], f["c55"],
                             f["bx"], f["bz"], fdc, *state, use_reentrant=True)
        elif fd_order != 4:
            out = checkpoint(_steps_abl6, free_surface, nx, nz, dx, dz, nabc, fso, sx, sz, dt, chunk, MT, rx, rz,
                             dmp, f["c11"], f["c13"], f["c15"], f["c33"], f["c35"], f["c55"],
                             f["bx"], f["bz"], fdc, *state, use_reentrant=True)
        elif pml:
            out = checkpoint(_steps_tuple, free_surface, nx, nz, dx, dz, nabc, sx, sz, dt, chunk, MT, rx, rz,
                             f["bcx"], f["bcz"], f["c11"], f["c13"], f["c15"], f["c33"], f["c35"], f["c55"],
                             f["bx"], f["bz"], fdc, *state, use_reentrant=True)
        else:
            out = checkpoint(_steps_abl_tuple, free_surface, nx, nz, dx, dz, nabc, sx, sz, dt, chunk, MT, rx, rz,
                             dmp, f["c11"], f["c13"], f["c15"], f["c33"], f["c35"], f["c55"],
                             f["bx"], f["bz"], fdc, *state, use_reentrant=True)
        state = list(out[:nstate])
        n = chunk.shape[-1]
        for k in range(5):
            recs[k][:, t0:t0 + n] = out[nstate + k]
            ens[k] += out[nstate + 5 + k].detach()
        t0 += n
    names = ("txx", "tzz", "txz", "vx", "vz")
    res = {n_: r for n_, r in zip(names, recs)}
    res.update({"forward_wavefield_" + n_: e for n_, e in zip(names, ens)})
    return res
