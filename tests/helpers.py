"""Shared drivers for the parity tests: run the oracle or the HIP path on a golden case."""
import numpy as np
import torch

from conftest import golden


def acoustic_inputs(name):
    g = golden(name)
    meta = dict(nz=int(g["nz"]), nx=int(g["nx"]), nabc=int(g["nabc"]), dx=float(g["dx"]),
                dt=float(g["dt"]), nt=int(g["nt"]), fs=bool(g["free_surface"]))
    return g, meta


def run_acoustic_oracle(g, m, seg=1, dtype=torch.float32, want_grad=True, weights=None):
    from oracle import acoustic as oa
    v = torch.tensor(g["vp"], dtype=dtype, requires_grad=want_grad)
    r = torch.tensor(g["rho"], dtype=dtype, requires_grad=want_grad)
    w = torch.tensor(g["wavelet"], dtype=dtype, requires_grad=want_grad)
    res = oa.forward_kernel(m["nx"], m["nz"], m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                            torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), len(g["src_x"]), w,
                            torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]), len(g["rcv_x"]),
                            torch.tensor(g["damp"], dtype=dtype), v, r, checkpoint_segments=seg,
                            dtype=dtype)
    out = {k: res[k].detach().numpy() for k in ("p", "u", "w")}
    out.update({"fw_" + k: res["forward_wavefield_" + k].numpy() for k in "puw"})
    if want_grad:
        wts = g["weights"] if weights is None else weights
        loss = sum((res[k] * torch.tensor(wts[i], dtype=dtype)).sum() for i, k in enumerate("puw"))
        loss.backward()
        out.update(grad_vp=v.grad.numpy(), grad_rho=r.grad.numpy(), grad_wavelet=w.grad.numpy())
    return out


def elastic_inputs(name):
    g = golden(name)
    meta = dict(nz=int(g["nz"]), nx=int(g["nx"]), nabc=int(g["nabc"]), dx=float(g["dx"]),
                dt=float(g["dt"]), nt=int(g["nt"]), fs=bool(g["free_surface"]))
    return g, meta


ELASTIC_LEAVES = ("c11", "c13", "c33", "c55", "bx", "bz")
ELASTIC_NAMES = ("txx", "tzz", "txz", "vx", "vz")


def elastic_cc(leaves, nz, nx):
    z = torch.zeros((nz, nx), dtype=leaves["c11"].dtype, device=leaves["c11"].device)
    return [leaves["c11"], z, leaves["c13"], z, z, z, z, z, z, z, z, leaves["c33"], z, z, z, z, z, z,
            leaves["c55"], z, z]


def run_elastic_oracle(g, m, seg=1, want_grad=True):
    from oracle import elastic as oe
    leaves = {k: torch.tensor(g[k], requires_grad=want_grad) for k in ELASTIC_LEAVES}
    nz, nx = m["nz"], m["nx"]
    z = torch.zeros((nz, nx))
    w = torch.tensor(g["wavelet"], requires_grad=want_grad)
    res = oe.forward_kernel(nx, nz, m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                            torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), len(g["src_x"]), w,
                            torch.tensor(g["mt"]), torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]),
                            len(g["rcv_x"]), "PML", torch.tensor(g["bcx"], dtype=torch.float32),
                            torch.tensor(g["bcz"], dtype=torch.float32), None, z, z,
                            leaves["bx"], leaves["bz"], elastic_cc(leaves, nz, nx), n_segments=seg)
    out = {k: res[k].detach().numpy() for k in ELASTIC_NAMES}
    out.update({"fw_" + k: res["forward_wavefield_" + k].numpy() for k in ELASTIC_NAMES})
    if want_grad:
        loss = sum((res[k] * torch.tensor(g["weights"][i])).sum() for i, k in enumerate(ELASTIC_NAMES))
        loss.backward()
        out.update({"grad_" + k: v.grad.numpy() for k, v in leaves.items()})
        out["grad_wavelet"] = w.grad.numpy()
    return out
