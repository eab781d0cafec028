"""Writes the golden vectors in this directory from the reference itself.

Run in the development container only (the reference tree is not present where
the tests run):

    python tests/golden/make_golden.py [case ...]

It loads three self-contained reference files by path -- they import only
torch/numpy/matplotlib, so no stubs are needed -- and records, on tiny grids,
the inputs and outputs of:

* ADFWI/propagator/acoustic_kernels.py  forward_kernel (:179-301): records
  p/u/w, the three energy maps, and autograd gradients of a seeded linear
  functional of all records with respect to vp, rho and the wavelet, for
  free surface off/on and checkpoint_segments 1 and 3;
* ADFWI/propagator/elastic_kernels.py   forward_kernel (:917-1037, 4th order):
  five records, five energy maps and gradients with respect to C11, C13, C33,
  C55, bx, bz and the wavelet, isotropic with free surface and VTI without,
  segments 1 and 3, for the split-field PML branch and for the damping (ABL)
  branch with Cerjan (free surface) and sin^2 (no free surface) tapers; the
  three-term schemes of fd_order 6, 8 and 10 (step_forward_PML_6order,
  :426-584, and step_forward_ABL_6order, :781-912); DiffCoef(2, 's'), DiffCoef(3, 's') and DiffCoef(4, 's') (:20-58);
* ADFWI/propagator/boundary_condition.py bc_pml, bc_pml_xz, bc_sincos,
  bc_gerjan (:15-120).

These functions' quirks (energy-map accumulation, free-surface rows, padded
shape offsets, fp32 association) are not fixed by any closed form or invariant,
which is why the vectors come from the reference.  Only data is written.
"""
import importlib.util
import os
import sys

import numpy as np
import torch

REF = "/root/reference/ADFWI/propagator"
OUT = os.path.dirname(os.path.abspath(__file__))


def _load(name, fname):
    spec = importlib.util.spec_from_file_location(name, os.path.join(REF, fname))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def ricker(nt, dt, f0):
    t = np.arange(nt) * dt
    t0 = 1.2 / f0
    tau = (np.pi * f0) ** 2
    return (1 - 2 * tau * (t - t0) ** 2) * np.exp(-tau * (t - t0) ** 2)


def smooth_field(rng, nz, nx, lo, hi, passes=4):
    a = rng.standard_normal((nz, nx))
    for _ in range(passes):
        a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
    a = (a - a.min()) / (a.max() - a.min())
    return lo + (hi - lo) * a


def acoustic_case(ak, bc, free_surface, seed):
    rng = np.random.default_rng(seed)
    nz, nx, nabc, dx, dt, nt = 16, 24, 6, 10.0, 1e-3, 120
    vp = smooth_field(rng, nz, nx, 1500.0, 3000.0).astype(np.float32)
    rho = smooth_field(rng, nz, nx, 1000.0, 2500.0).astype(np.float32)
    src_x = np.array([5, 17]); src_z = np.array([2, 3]) if not free_surface else np.array([1, 3])
    rcv_x = np.arange(nx); rcv_z = np.ones(nx, dtype=np.int64)
    wav = np.stack([ricker(nt, dt, 25.0), 0.7 * ricker(nt, dt, 30.0)]).astype(np.float32)
    damp = bc.bc_pml(nx, nz, dx, dx, pml=nabc, vmax=float(vp.max()), free_surface=False)
    wts = rng.standard_normal((3, 2, nt, nx)).astype(np.float32)
    out = dict(nz=nz, nx=nx, nabc=nabc, dx=dx, dt=dt, nt=nt, free_surface=int(free_surface),
               vp=vp, rho=rho, src_x=src_x, src_z=src_z, rcv_x=rcv_x, rcv_z=rcv_z, wavelet=wav,
               damp=damp, weights=wts)
    for seg in (1, 3):
        v = torch.tensor(vp, requires_grad=True)
        r = torch.tensor(rho, requires_grad=True)
        w = torch.tensor(wav, requires_grad=True)
        res = ak.forward_kernel(nx, nz, dx, dx, nt, dt, nabc, free_surface,
                                torch.tensor(src_x), torch.tensor(src_z), 2, w,
                                torch.tensor(rcv_x), torch.tensor(rcv_z), nx,
                                torch.tensor(damp, dtype=torch.float32), v, r,
                                checkpoint_segments=seg)
        loss = sum((res[k] * torch.tensor(wts[i])).sum() for i, k in enumerate("puw"))
        loss.backward()
        for k in ("p", "u", "w"):
            out[f"seg{seg}_rec_{k}"] = res[k].detach().numpy()
            out[f"seg{seg}_fw_{k}"] = res["forward_wavefield_" + k].numpy()
        out[f"seg{seg}_grad_vp"] = v.grad.numpy()
        out[f"seg{seg}_grad_rho"] = r.grad.numpy()
        out[f"seg{seg}_grad_wavelet"] = w.grad.numpy()
    return out


def elastic_case(ek, bc, free_surface, vti, seed, abc="PML", order=4):
    rng = np.random.default_rng(seed)
    nz, nx, nabc, dx, dt, nt = 16, 24, 6, 5.0, 4e-4, 120
    vp = smooth_field(rng, nz, nx, 1800.0, 3000.0)
    vs = vp / 1.73 * smooth_field(rng, nz, nx, 0.95, 1.05)
    rho = smooth_field(rng, nz, nx, 1800.0, 2400.0)
    eps = smooth_field(rng, nz, nx, 0.0, 0.25) if vti else np.zeros((nz, nx))
    delta = smooth_field(rng, nz, nx, -0.05, 0.15) if vti else np.zeros((nz, nx))
    c33 = vp ** 2 * rho
    c44 = vs ** 2 * rho
    c11 = c33 * (1 + 2 * eps)
    c13 = np.sqrt(2 * c33 * (c33 - c44) * delta + (c33 - c44) ** 2) - c44
    b = 1 / rho
    bx = 0.5 * (b[:, :-1] + b[:, 1:])
    bz = 0.5 * (b[:-1, :] + b[1:, :])
    c55 = 0.2 * (c44[1:-1, 1:-1] + c44[2:, 1:-1] + c44[1:-1, 2:] + c44[2:, 1:-1] + c44[2:, 2:])
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    c11, c13, c33, c55, bx, bz = map(f32, (c11, c13, c33, c55, bx, bz))
    src_x = np.array([6, 15]); src_z = np.array([3, 5])
    rcv_x = np.arange(nx); rcv_z = np.full(nx, 1, dtype=np.int64)
    wav = np.stack([1e6 * ricker(nt, dt, 60.0), 0.8e6 * ricker(nt, dt, 50.0)]).astype(np.float32)
    mt = np.stack([np.eye(3), np.array([[1.0, 0, 0.5], [0, 1, 0], [0.5, 0, -0.7]])]).astype(np.float32)
    bcx, bcz = bc.bc_pml_xz(nx, nz, dx, dx, pml=nabc, vmax=float(vp.max()), free_surface=free_surface)
    if abc == "gerjan":
        damp = bc.bc_gerjan(nx, nz, dx, dx, pml=nabc, alpha=0.0053, free_surface=free_surface)
    elif abc != "PML":
        damp = bc.bc_sincos(nx, nz, dx, dx, pml=nabc, free_surface=free_surface)
    else:
        damp = np.zeros((1, 1))
    wts = rng.standard_normal((5, 2, nt, nx)).astype(np.float32)
    out = dict(nz=nz, nx=nx, nabc=nabc, dx=dx, dt=dt, nt=nt, free_surface=int(free_surface), fd_order=order,
               c11=c11, c13=c13, c33=c33, c55=c55, bx=bx, bz=bz, src_x=src_x, src_z=src_z,
               rcv_x=rcv_x, rcv_z=rcv_z, wavelet=wav, mt=mt, bcx=bcx, bcz=bcz, weights=wts,
               vp=vp, vs=vs, rho=rho, eps=eps, delta=delta, abc=abc, damp=damp)
    names = ("txx", "tzz", "txz", "vx", "vz")
    for seg in (1, 3):
        leaves = {k: torch.tensor(out[k], requires_grad=True) for k in ("c11", "c13", "c33", "c55", "bx", "bz")}
        z = torch.zeros((nz, nx))
        CC = [leaves["c11"], z, leaves["c13"], z, z, z, z, z, z, z, z, leaves["c33"], z, z, z, z, z, z,
              leaves["c55"], z, z]
        w = torch.tensor(wav, requires_grad=True)
        res = ek.forward_kernel(nx, nz, dx, dx, nt, dt, nabc, free_surface,
                                torch.tensor(src_x), torch.tensor(src_z), 2, w, torch.tensor(mt),
                                torch.tensor(rcv_x), torch.tensor(rcv_z), nx,
                                abc, torch.tensor(bcx, dtype=torch.float32),
                                torch.tensor(bcz, dtype=torch.float32),
                                None if abc == "PML" else torch.tensor(damp, dtype=torch.float32),
                                z, z, leaves["bx"], leaves["bz"], CC, fd_order=order, n_segments=seg)
        loss = sum((res[k] * torch.tensor(wts[i])).sum() for i, k in enumerate(names))
        loss.backward()
        for k in names:
            out[f"seg{seg}_rec_{k}"] = res[k].detach().numpy()
            out[f"seg{seg}_fw_{k}"] = res["forward_wavefield_" + k].numpy()
        for k, v in leaves.items():
            out[f"seg{seg}_grad_{k}"] = v.grad.numpy()
        out[f"seg{seg}_grad_wavelet"] = w.grad.numpy()
    return out


def boundary_case(bc):
    out = {}
    out["pml_nofs"] = bc.bc_pml(24, 16, 10.0, 10.0, pml=6, vmax=3000.0, free_surface=False)
    out["pml_fs"] = bc.bc_pml(24, 16, 10.0, 10.0, pml=6, vmax=3000.0, free_surface=True)
    for fs in (False, True):
        bcx, bcz = bc.bc_pml_xz(24, 16, 5.0, 5.0, pml=6, vmax=3000.0, free_surface=fs)
        out[f"pmlxz_fs{int(fs)}_x"] = bcx
        out[f"pmlxz_fs{int(fs)}_z"] = bcz
        out[f"sincos_fs{int(fs)}"] = bc.bc_sincos(24, 16, 10.0, 10.0, pml=6, free_surface=fs)
        out[f"gerjan_fs{int(fs)}"] = bc.bc_gerjan(24, 16, 10.0, 10.0, pml=6, alpha=0.0053, free_surface=fs)
    return out


def main():
    torch.manual_seed(0)
    torch.set_num_threads(1)
    ak = _load("_ref_acoustic_kernels", "acoustic_kernels.py")
    ek = _load("_ref_elastic_kernels", "elastic_kernels.py")
    bc = _load("_ref_boundary_condition", "boundary_condition.py")
    cases = {
        "acoustic_nofs": lambda: acoustic_case(ak, bc, False, 11),
        "acoustic_fs": lambda: acoustic_case(ak, bc, True, 12),
        "elastic_iso_fs": lambda: elastic_case(ek, bc, True, False, 21),
        "elastic_vti_nofs": lambda: elastic_case(ek, bc, False, True, 22),
        "elastic_abl_gerjan_fs": lambda: elastic_case(ek, bc, True, False, 23, abc="gerjan"),
        "elastic_abl_sincos_nofs": lambda: elastic_case(ek, bc, False, True, 24, abc="sincos"),
        "elastic_fd6_iso_fs": lambda: elastic_case(ek, bc, True, False, 25, order=6),
        "elastic_fd6_vti_nofs": lambda: elastic_case(ek, bc, False, True, 26, order=6),
        "elastic_fd8_iso_fs": lambda: elastic_case(ek, bc, True, False, 27, order=8),
        "elastic_fd10_vti_nofs": lambda: elastic_case(ek, bc, False, True, 28, order=10),
        "elastic_fd6_abl_gerjan_fs": lambda: elastic_case(ek, bc, True, False, 29, abc="gerjan", order=6),
        "elastic_fd8_abl_sincos_nofs": lambda: elastic_case(ek, bc, False, True, 30, abc="sincos", order=8),
        "boundary": lambda: boundary_case(bc),
        "diffcoef": lambda: dict(fdc=ek.DiffCoef(2, "s").numpy()),
        "diffcoef6": lambda: dict(fdc6=ek.DiffCoef(3, "s").numpy(), fdc8=ek.DiffCoef(4, "s").numpy(),
                                  fdc10=ek.DiffCoef(5, "s").numpy()),
    }
    names = sys.argv[1:] or list(cases)   # write only the named cases when given
    for name in names:
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **cases[name]())
    print("golden vectors written to", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
