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
                dt=float(g["dt"]), nt=int(g["nt"]), fs=bool(g["free_surface"]),
                fd=int(g["fd_order"]) if "fd_order" in g else 4)
    return g, meta


ELASTIC_LEAVES = ("c11", "c13", "c33", "c55", "bx", "bz")
ELASTIC_NAMES = ("txx", "tzz", "txz", "vx", "vz")


def elastic_cc(leaves, nz, nx):
    z = torch.zeros((nz, nx), dtype=leaves["c11"].dtype, device=leaves["c11"].device)
    return [leaves["c11"], z, leaves["c13"], z, z, z, z, z, z, z, z, leaves["c33"], z, z, z, z, z, z,
            leaves["c55"], z, z]


def elastic_abc(g):
    """abc_type of a golden elastic case ('PML' unless the case stores another)."""
    return str(g["abc"]) if "abc" in g else "PML"


def elastic_damp(g):
    return None if elastic_abc(g) == "PML" else torch.tensor(g["damp"], dtype=torch.float32)


def run_elastic_oracle(g, m, seg=1, want_grad=True):
    from oracle import elastic as oe
    leaves = {k: torch.tensor(g[k], requires_grad=want_grad) for k in ELASTIC_LEAVES}
    nz, nx = m["nz"], m["nx"]
    z = torch.zeros((nz, nx))
    w = torch.tensor(g["wavelet"], requires_grad=want_grad)
    res = oe.forward_kernel(nx, nz, m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                            torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), len(g["src_x"]), w,
                            torch.tensor(g["mt"]), torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]),
                            len(g["rcv_x"]), elastic_abc(g), torch.tensor(g["bcx"], dtype=torch.float32),
                            torch.tensor(g["bcz"], dtype=torch.float32), elastic_damp(g), z, z,
                            leaves["bx"], leaves["bz"], elastic_cc(leaves, nz, nx), n_segments=seg,
                            fd_order=int(g["fd_order"]) if "fd_order" in g else 4)
    out = {k: res[k].detach().numpy() for k in ELASTIC_NAMES}
    out.update({"fw_" + k: res["forward_wavefield_" + k].numpy() for k in ELASTIC_NAMES})
    if want_grad:
        loss = sum((res[k] * torch.tensor(g["weights"][i])).sum() for i, k in enumerate(ELASTIC_NAMES))
        loss.backward()
        out.update({"grad_" + k: v.grad.numpy() for k, v in leaves.items()})
        out["grad_wavelet"] = w.grad.numpy()
    return out


def _smooth(rng, nz, nx, lo, hi, passes=4):
    a = rng.standard_normal((nz, nx))
    for _ in range(passes):
        a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
    a = (a - a.min()) / (a.max() - a.min())
    return lo + (hi - lo) * a


def elastic_random_case(nz, nx, nabc, nt, nshots, free_surface, seed, aniso=False, dx=5.0, dt=4e-4, abc="PML",
                        fd_order=4):
    """A golden-shaped elastic case on a larger grid (no reference outputs): smooth fields,
    staggered shapes as the reference's models build them, shots spread over the grid,
    receivers along one row.  aniso adds small C15/C35 fields."""
    from oracle import boundary as ob
    rng = np.random.default_rng(seed)
    vp = _smooth(rng, nz, nx, 1800.0, 3000.0)
    vs = vp / 1.73
    rho = _smooth(rng, nz, nx, 1800.0, 2400.0)
    c33 = vp ** 2 * rho
    c44 = vs ** 2 * rho
    c11 = c33 * 1.1
    c13 = c33 - 2 * c44
    b = 1 / rho
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    g = dict(c11=f32(c11), c13=f32(c13), c33=f32(c33),
             c55=f32(0.2 * (c44[1:-1, 1:-1] + c44[2:, 1:-1] + c44[1:-1, 2:] + c44[2:, 1:-1] + c44[2:, 2:])),
             bx=f32(0.5 * (b[:, :-1] + b[:, 1:])), bz=f32(0.5 * (b[:-1, :] + b[1:, :])))
    if aniso:
        g["c15"] = f32(0.05 * c33 * _smooth(rng, nz, nx, -1.0, 1.0))
        g["c35"] = f32(0.04 * c33 * _smooth(rng, nz, nx, -1.0, 1.0))
    g["src_x"] = np.linspace(3, nx - 4, nshots).astype(np.int64)
    g["src_z"] = (2 + np.arange(nshots) % max(1, nz - 4)).astype(np.int64)
    g["rcv_x"] = np.arange(0, nx, 2, dtype=np.int64)
    g["rcv_z"] = np.full(len(g["rcv_x"]), 1, dtype=np.int64)
    t = np.arange(nt) * dt
    f0 = 60.0
    tau = (np.pi * f0) ** 2
    ric = (1 - 2 * tau * (t - 1.2 / f0) ** 2) * np.exp(-tau * (t - 1.2 / f0) ** 2)
    g["wavelet"] = f32(np.stack([1e6 * (1 + 0.1 * s) * ric for s in range(nshots)]))
    mt = np.zeros((nshots, 3, 3), dtype=np.float32)
    for s in range(nshots):
        mt[s] = np.array([[1.0, 0, 0.3 * (s % 3)], [0, 1, 0], [0.3 * (s % 3), 0, 1.0 - 0.5 * (s % 2)]])
    g["mt"] = mt
    g["bcx"], g["bcz"] = ob.bc_pml_xz(nx, nz, dx, dx, pml=nabc, vmax=float(vp.max()), free_surface=free_surface)
    if abc != "PML":
        g["abc"] = np.array(abc)
        g["damp"] = (ob.bc_gerjan(nx, nz, dx, dx, nabc, alpha=0.0053, free_surface=free_surface) if abc == "gerjan"
                     else ob.bc_sincos(nx, nz, dx, dx, nabc, free_surface=free_surface))
    g["weights"] = rng.standard_normal((5, nshots, nt, len(g["rcv_x"]))).astype(np.float32)
    meta = dict(nz=nz, nx=nx, nabc=nabc, dx=dx, dt=dt, nt=nt, fs=bool(free_surface), fd=int(fd_order))
    return g, meta


def elastic_cc_general(g, nz, nx, dtype=torch.float32):
    """21-entry CC list of a golden-shaped case (C15/C35 when present)."""
    t = lambda k: torch.tensor(g[k], dtype=dtype) if k in g else torch.zeros((nz, nx), dtype=dtype)
    z = torch.zeros((nz, nx), dtype=dtype)
    return [t("c11"), z, t("c13"), z, t("c15"), z, z, z, z, z, z, t("c33"), z, t("c35"), z, z, z, z,
            t("c55"), z, z]
