"""End-to-end gradients on the GPU: the drop-in propagators driven through the model classes
(vp / vs / rho / Thomsen parameters, as the FWI drivers use them) against the oracle's autograd,
and central finite differences along seeded random directions.  Tolerances: rel-L2 <= 1e-4
against the oracle; finite differences to 5e-3 (fp32 records, double-accumulated loss)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _smooth(rng, nz, nx, lo, hi, passes=6):
    a = rng.standard_normal((nz, nx))
    for _ in range(passes):
        a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
    a = (a - a.min()) / (a.max() - a.min())
    return (lo + (hi - lo) * a).astype(np.float32)


def _ricker(nt, dt, f0):
    t = np.arange(nt) * dt - 1.2 / f0
    tau = (np.pi * f0) ** 2
    return (1 - 2 * tau * t ** 2) * np.exp(-tau * t ** 2)


def _elastic_case(vti, seed):
    rng = np.random.default_rng(seed)
    nz, nx, nabc, nt, dt, dx = 30, 64, 8, 90, 4e-4, 5.0
    vp = _smooth(rng, nz, nx, 1900.0, 2800.0)
    vs = (vp / 1.73 * _smooth(rng, nz, nx, 0.95, 1.05)).astype(np.float32)
    rho = _smooth(rng, nz, nx, 1900.0, 2300.0)
    eps = _smooth(rng, nz, nx, 0.0, 0.2) if vti else np.zeros((nz, nx), np.float32)
    delta = _smooth(rng, nz, nx, -0.05, 0.1) if vti else np.zeros((nz, nx), np.float32)
    return dict(nz=nz, nx=nx, nabc=nabc, nt=nt, dt=dt, dx=dx, fs=not vti, vp=vp, vs=vs, rho=rho, eps=eps,
                delta=delta, src_x=np.array([10, 33, 55]), src_z=np.array([3, 6, 2]), rcv_x=np.arange(0, nx, 2),
                rcv_z=np.full(nx // 2, 1), wav=(1e6 * _ricker(nt, dt, 50.0)).astype(np.float32),
                wts=rng.standard_normal((2, 3, nt, nx // 2)).astype(np.float32))


def _elastic_propagator(c, vti, grad=True):
    from adfwi_amd.model import AnisotropicElasticModel, IsotropicElasticModel
    from adfwi_amd.propagator import ElasticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    src = Source(c["nt"], c["dt"], 50.0)
    src.add_sources(c["src_x"], c["src_z"], c["wav"])
    rcv = Receiver(c["nt"], c["dt"])
    rcv.add_receivers(c["rcv_x"], c["rcv_z"], "vz")
    kw = dict(vp=c["vp"], vs=c["vs"], rho=c["rho"], vp_grad=grad, vs_grad=grad, rho_grad=grad,
              free_surface=c["fs"], nabc=c["nabc"], auto_update_rho=False, device=DEV)
    if vti:
        model = AnisotropicElasticModel(0, 0, c["nx"], c["nz"], c["dx"], c["dx"], eps=c["eps"], delta=c["delta"],
                                        gamma=np.zeros_like(c["eps"]), eps_grad=grad, delta_grad=grad, **kw)
    else:
        model = IsotropicElasticModel(0, 0, c["nx"], c["nz"], c["dx"], c["dx"], **kw)
    return ElasticPropagator(model, Survey(src, rcv), device=DEV), model


def _elastic_loss(res, wts):
    return sum((res[k].double() * torch.as_tensor(wts[i], device=res[k].device).double()).sum()
               for i, k in enumerate(("vx", "vz")))


@pytest.mark.parametrize("vti", [False, True])
def test_elastic_model_gradients_match_oracle(vti):
    """d loss / d (vp, vs, rho[, eps, delta]) through the model parameterisation, the padding and
    the hand-written adjoint, against the oracle's autograd through the same parameterisation."""
    from oracle import boundary as ob
    from oracle import elastic as oe
    from oracle import model as om
    c = _elastic_case(vti, seed=31 + int(vti))
    prop, model = _elastic_propagator(c, vti)
    _elastic_loss(prop.forward(), c["wts"]).backward()
    names = ["vp", "vs", "rho"] + (["eps", "delta"] if vti else [])
    got = {n: getattr(model, n).grad.cpu().numpy() for n in names}
    leaves = {n: torch.tensor(c[n], requires_grad=n in names) for n in ("vp", "vs", "rho", "eps", "delta")}
    nz, nx = c["nz"], c["nx"]
    CC, bx, bz, lamu, lam = om.elastic_cc(leaves["vp"], leaves["vs"], leaves["rho"], leaves["eps"],
                                          leaves["delta"], torch.zeros((nz, nx)), nx, nz)
    bcx, bcz = ob.bc_pml_xz(nx, nz, c["dx"], c["dx"], c["nabc"], float(c["vp"].max()), c["fs"])
    S = len(c["src_x"])
    res = oe.forward_kernel(nx, nz, c["dx"], c["dx"], c["nt"], c["dt"], c["nabc"], c["fs"],
                            torch.tensor(c["src_x"]), torch.tensor(c["src_z"]), S,
                            torch.tensor(np.tile(c["wav"], (S, 1))), torch.eye(3).repeat(S, 1, 1),
                            torch.tensor(c["rcv_x"]), torch.tensor(c["rcv_z"]), len(c["rcv_x"]), "PML",
                            torch.tensor(bcx, dtype=torch.float32), torch.tensor(bcz, dtype=torch.float32), None,
                            lamu, lam, bx, bz, CC)
    _elastic_loss(res, c["wts"]).backward()
    for n in names:
        ref = leaves[n].grad.numpy()
        assert rel_l2(got[n], ref) <= 1e-4, (n, rel_l2(got[n], ref))


@pytest.mark.parametrize("fd_order", [4, 6])
def test_elastic_finite_difference_along_random_direction(fd_order):
    """<grad, d> against the central difference of the loss along a seeded smooth direction d in
    (vp, vs, rho), with the model classes' parameterisation recomputed for each perturbation, for
    the 4th-order and the three-term (6th-order) stencil."""
    c = _elastic_case(False, seed=41)
    rng = np.random.default_rng(5)
    dirs = {n: _smooth(rng, c["nz"], c["nx"], -1.0, 1.0) * scale for n, scale in
            (("vp", 20.0), ("vs", 12.0), ("rho", 15.0))}
    prop, model = _elastic_propagator(c, False)
    _elastic_loss(prop.forward(fd_order=fd_order), c["wts"]).backward()
    dd = sum(float((getattr(model, n).grad.double().cpu() * torch.tensor(d).double()).sum()) for n, d in dirs.items())
    vals = []
    for sgn in (1.0, -1.0):
        cc = dict(c)
        for n, d in dirs.items():
            cc[n] = (c[n] + sgn * d).astype(np.float32)
        p, _ = _elastic_propagator(cc, False, grad=False)
        with torch.no_grad():
            vals.append(float(_elastic_loss(p.forward(fd_order=fd_order), c["wts"])))
    fd = (vals[0] - vals[1]) / 2.0
    assert abs(fd - dd) <= 5e-3 * abs(dd), (fd, dd)


def test_acoustic_finite_difference_along_random_direction():
    """<d loss / d vp, d> against the central difference along a seeded smooth direction, rho held
    fixed (the drop-in forward_kernel with vp and rho as inputs)."""
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.propagator.boundary_condition import bc_pml
    rng = np.random.default_rng(12)
    nz, nx, nabc, nt, dt, dx = 40, 90, 10, 300, 1e-3, 10.0
    vp = _smooth(rng, nz, nx, 1800.0, 3200.0)
    rho = _smooth(rng, nz, nx, 1900.0, 2400.0)
    damp = torch.tensor(bc_pml(nx, nz, dx, dx, nabc, 3300.0, False), dtype=torch.float32, device=DEV)
    sx, sz = torch.tensor([15, 45, 75]), torch.tensor([2, 2, 2])
    rx, rz = torch.arange(nx), torch.full((nx,), 3)
    wav = torch.tensor(np.tile(_ricker(nt, dt, 15.0), (3, 1)), dtype=torch.float32, device=DEV)
    wts = torch.tensor(rng.standard_normal((3, nt, nx)), dtype=torch.float64, device=DEV)
    d = torch.tensor(_smooth(rng, nz, nx, -1.0, 1.0) * 10.0, device=DEV)

    def loss(v, r):
        res = ak.forward_kernel(nx, nz, dx, dx, nt, dt, nabc, False, sx, sz, 3, wav, rx, rz, nx, damp, v, r,
                                device=DEV)
        return (res["p"].double() * wts).sum()
    v = torch.tensor(vp, device=DEV, requires_grad=True)
    r = torch.tensor(rho, device=DEV)
    loss(v, r).backward()
    dd = float((v.grad.double() * d.double()).sum())
    with torch.no_grad():
        fd = (float(loss(v.detach() + d, r)) - float(loss(v.detach() - d, r))) / 2.0
    assert abs(fd - dd) <= 5e-3 * abs(dd), (fd, dd)
