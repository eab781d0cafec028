"""Vertical-force sources on the elastic propagator (an extension: the reference has moment-
tensor sources only, so this path is pinned by properties rather than reference vectors).

* Mirror symmetries in a homogeneous medium before the waves reach a boundary: a vertical
  force at a vz node gives vz symmetric and vx antisymmetric under reflection about the force
  (in z about the vz node, index i <-> 2 zs - i; in x about the source column, vz index
  j <-> 2 xs - j, vx index j <-> 2 xs - 1 - j).  An explosive source at the same indices sits on
  a stress node and does not give the z symmetry of vz, so the test also pins where and what
  is injected.
* Linearity in the wavelet (scaling and superposition, to rounding).
* The wavelet gradient is the adjoint of the wavelet -> records map (dot-product test), in the
  PML and the damping scheme, also for a shot that has both moment-tensor and force weights.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _ricker(nt, dt, f0):
    t = np.arange(nt) * dt - 1.2 / f0
    a = (np.pi * f0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _run(src_type, rcv, nt=70, abc="PML", wav=None, n=61, nabc=10):
    from adfwi_amd.model import IsotropicElasticModel
    from adfwi_amd.propagator import ElasticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    dt = 4e-4
    wav = _ricker(nt, dt, 40.0) if wav is None else wav
    src = Source(nt, dt, 40.0)
    src.add_sources([n // 2], [n // 2], wav, src_type)
    r = Receiver(nt, dt)
    r.add_receivers(np.array([p[1] for p in rcv]), np.array([p[0] for p in rcv]), "vz")
    model = IsotropicElasticModel(0, 0, n, n, 5.0, 5.0, vp=np.full((n, n), 2500.0, np.float32),
                                  vs=np.full((n, n), 1450.0, np.float32), rho=np.full((n, n), 2100.0, np.float32),
                                  free_surface=False, nabc=nabc, abc_type=abc, auto_update_rho=False, device=DEV)
    with torch.no_grad():
        out = ElasticPropagator(model, Survey(src, r), device=DEV).forward()
    return {k: out[k][0].cpu().numpy() for k in ("vx", "vz")}


def test_vertical_force_mirror_symmetries():
    n = 61
    c = n // 2
    ms = range(1, 9)
    # z mirror at the source column, x mirror at the source row (vx pairs use 2 xs - 1 - j)
    rcv = ([(c + m, c) for m in ms] + [(c - m, c) for m in ms] + [(c, c + m) for m in ms] + [(c, c - m) for m in ms]
           + [(c, c + m) for m in ms] + [(c, c - 1 - m) for m in ms])
    k = len(list(ms))
    rec = _run("vf", rcv)
    vz, vx = rec["vz"], rec["vx"]
    scale = np.abs(vz).max()
    assert scale > 0
    up, down = vz[:, 0:k], vz[:, k:2 * k]
    right, left = vz[:, 2 * k:3 * k], vz[:, 3 * k:4 * k]
    assert np.allclose(up, down, rtol=0, atol=1e-5 * scale), np.abs(up - down).max() / scale
    assert np.allclose(right, left, rtol=0, atol=1e-5 * scale)
    vxr, vxl = vx[:, 4 * k:5 * k], vx[:, 5 * k:6 * k]
    assert np.abs(vxr).max() > 1e-3 * np.abs(vx).max() + 0
    assert np.allclose(vxr, -vxl, rtol=0, atol=1e-5 * np.abs(vx).max())
    # an explosive source at the same indices breaks the up-down symmetry of vz about that node
    ex = _run("mt", rcv)["vz"]
    sc = np.abs(ex).max()
    assert np.abs(ex[:, 0:k] - ex[:, k:2 * k]).max() > 1e-2 * sc


def test_vertical_force_is_linear_in_the_wavelet():
    rcv = [(5, j) for j in range(5, 56, 5)]
    w1, w2 = _ricker(70, 4e-4, 40.0), _ricker(70, 4e-4, 25.0)
    a, b = _run("vf", rcv, wav=w1), _run("vf", rcv, wav=w2)
    two = _run("vf", rcv, wav=2 * w1)
    # doubling is exact except where a product rounds into the subnormal range (the leading edge)
    assert rel_l2(two["vz"], 2 * a["vz"]) < 1e-7 and rel_l2(two["vx"], 2 * a["vx"]) < 1e-7
    s = _run("vf", rcv, wav=w1 + w2)
    assert rel_l2(s["vz"], a["vz"] + b["vz"]) < 1e-5


@pytest.mark.parametrize("abc", ["PML", "gerjan"])
@pytest.mark.parametrize("mt_too", [False, True])
def test_wavelet_gradient_is_the_adjoint_with_a_force(abc, mt_too):
    """<F w, d> = <w, F^T d> for the wavelets of a force shot (and of a shot carrying both a
    moment tensor and a force weight, through forward_kernel's src_fz)."""
    from adfwi_amd.model import IsotropicElasticModel
    from adfwi_amd.propagator import boundary_condition as bcm
    from adfwi_amd.propagator.elastic_kernels import forward_kernel
    rng = np.random.default_rng(3)
    nz, nx, nabc, nt, dt, dx = 40, 50, 8, 80, 4e-4, 5.0
    vp = (2400 + 300 * rng.random((nz, nx))).astype(np.float32)
    model = IsotropicElasticModel(0, 0, nx, nz, dx, dx, vp=vp, vs=(vp / 1.7).astype(np.float32),
                                  rho=np.full((nz, nx), 2000.0, np.float32), free_surface=True, nabc=nabc,
                                  abc_type=abc, auto_update_rho=False, device=DEV)
    model.forward()
    S = 2
    sx, sz = torch.tensor([12, 35]), torch.tensor([6, 9])
    rx, rz = torch.arange(2, nx - 2, 3), torch.full((len(range(2, nx - 2, 3)),), 1)
    mt = torch.eye(3).repeat(S, 1, 1) * (1.0 if mt_too else 0.0)
    fz = torch.tensor([1.0, 0.7])
    if abc == "PML":
        bcx, bcz = bcm.bc_pml_xz(nx, nz, dx, dx, nabc, float(vp.max()), True)
        bcx, bcz, damp = torch.tensor(bcx, dtype=torch.float32), torch.tensor(bcz, dtype=torch.float32), None
    else:
        bcx = bcz = None
        damp = torch.tensor(bcm.bc_gerjan(nx, nz, dx, dx, nabc, 0.0053, True), dtype=torch.float32)
    d = {k: torch.tensor(rng.standard_normal((S, nt, len(rx))), dtype=torch.float64, device=DEV) for k in ("vx", "vz")}

    def records(w):
        return forward_kernel(nx, nz, dx, dx, nt, dt, nabc, True, sx, sz, S, w, mt, rx, rz, len(rx), abc, bcx, bcz,
                              damp, model.lamu, model.lam, model.bx, model.bz, model.CC, device=DEV, src_fz=fz)

    w = torch.tensor(rng.standard_normal((S, nt)).astype(np.float32), device=DEV, requires_grad=True)
    rec = records(w)
    lhs_loss = sum((rec[k].double() * d[k]).sum() for k in d)
    lhs_loss.backward()
    dw = torch.tensor(rng.standard_normal((S, nt)).astype(np.float32), device=DEV)
    with torch.no_grad():
        rec_dw = records(dw)
    lhs = float(sum((rec_dw[k].double() * d[k]).sum() for k in d))
    rhs = float((w.grad.double() * dw.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs)
