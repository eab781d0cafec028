"""Record and gradient post-processing on the GPU against the oracle: the fused trace-L2
kernel (misfit value rel <= 1e-6 and adjoint source rel-L2 <= 1e-5 against float64 autograd
of the reference expressions), the misfit driving the propagator's gradient, and the device
GradProcessor (float64 on both sides, rel <= 1e-10)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _records(seed, S=3, nt=240, R=37):
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 1, nt)[None, :, None]
    rec = (np.sin(40 * t + rng.uniform(0, 6, (S, 1, R))) * rng.uniform(0.5, 2, (S, 1, R))
           + 0.1 * rng.standard_normal((S, nt, R)))
    obs = np.sin(41 * t + rng.uniform(0, 6, (S, 1, R))) + 0.1 * rng.standard_normal((S, nt, R))
    return rec.astype(np.float32), obs.astype(np.float32)


@pytest.mark.parametrize("shape", [(3, 240, 37), (2, 5, 130), (1, 1001, 64)])
@pytest.mark.parametrize("normalize", [False, True])
def test_trace_l2_kernel_matches_oracle(normalize, shape):
    """Shapes: receivers fewer than, more than and exactly one 64-lane group; nt shorter than
    the number of time chunks and not a multiple of it."""
    from adfwi_amd.fwi import Misfit_waveform_L2, NormalizedWaveformL2
    from oracle import misfit as om
    S, nt, R = shape
    rec, obs = _records(5, S, nt, R)
    dt = 0.002
    obs_n = obs / np.abs(obs).max(axis=1, keepdims=True)
    a = torch.tensor(rec, device=DEV, requires_grad=True)
    b = torch.tensor(obs_n if normalize else obs, device=DEV, requires_grad=not normalize)
    loss = NormalizedWaveformL2(dt).forward(a, b) if normalize else Misfit_waveform_L2(dt).forward(a, b)
    loss.backward()
    ra = torch.tensor(rec, dtype=torch.float64, requires_grad=True)
    rb = torch.tensor(obs_n if normalize else obs, dtype=torch.float64, requires_grad=not normalize)
    ref = om.driver_misfit(ra, rb, dt, True) if normalize else om.waveform_l2(ra, rb, dt)
    ref.backward()
    assert abs(float(loss) - float(ref)) <= 1e-6 * abs(float(ref)), (float(loss), float(ref))
    assert rel_l2(a.grad.cpu().numpy(), ra.grad.numpy()) <= 1e-5
    if not normalize:
        assert rel_l2(b.grad.cpu().numpy(), rb.grad.numpy()) <= 1e-5


@pytest.mark.parametrize("shape", [(3, 240, 37), (1, 1001, 64)])
@pytest.mark.parametrize("normalize", [False, True])
def test_least_squares_mode_matches_oracle(normalize, shape):
    """ADFWI_L2_SQUARED: 1/2 sum (n - b)^2 dt and its adjoint sources against float64 autograd of
    the oracle expression, with and without the trace normalisation."""
    from adfwi_amd.fwi.misfit import NORMALIZE, SQUARED, _TraceL2
    from oracle import misfit as om
    S, nt, R = shape
    rec, obs = _records(6, S, nt, R)
    dt = 0.002
    a = torch.tensor(rec, device=DEV, requires_grad=True)
    b = torch.tensor(obs, device=DEV, requires_grad=True)
    loss = _TraceL2.apply(a, b, dt, SQUARED | (NORMALIZE if normalize else 0))
    loss.backward()
    ra = torch.tensor(rec, dtype=torch.float64, requires_grad=True)
    rb = torch.tensor(obs, dtype=torch.float64, requires_grad=True)
    ref = om.waveform_l2_squared(om.normalize_traces(ra) if normalize else ra, rb, dt)
    ref.backward()
    assert abs(float(loss) - float(ref)) <= 1e-6 * abs(float(ref)), (float(loss), float(ref))
    assert rel_l2(a.grad.cpu().numpy(), ra.grad.numpy()) <= 1e-5
    assert rel_l2(b.grad.cpu().numpy(), rb.grad.numpy()) <= 1e-5


def test_zero_residual_trace_nan_for_root_zero_for_least_squares():
    """A trace whose synthetic and observed samples are all equal (a far-offset trace that no
    arrival reaches) has a NaN adjoint source under the reference's per-trace square root, as
    torch autograd gives there, and a zero one under the least-squares objective; the other
    traces are unaffected."""
    from adfwi_amd.fwi import Misfit_waveform_L2, Misfit_waveform_L2_squared
    from oracle import misfit as om
    rec, obs = _records(7, 2, 120, 5)
    rec[1, :, 3] = 0.0
    obs[1, :, 3] = 0.0
    ref = torch.tensor(rec, dtype=torch.float64, requires_grad=True)
    om.waveform_l2(torch.tensor(obs, dtype=torch.float64), ref, 0.01).backward()
    assert torch.isnan(ref.grad[1, :, 3]).all()
    for cls, nan in ((Misfit_waveform_L2, True), (Misfit_waveform_L2_squared, False)):
        a = torch.tensor(rec, device=DEV, requires_grad=True)
        cls(0.01).forward(torch.tensor(obs, device=DEV), a).backward()
        g = a.grad.cpu()
        assert bool(torch.isnan(g[1, :, 3]).all()) == nan, cls.__name__
        assert torch.isfinite(g[0]).all() and torch.isfinite(g[1, :, :3]).all()
        if not nan:
            assert (g[1, :, 3] == 0).all()


def test_normalised_misfit_drives_the_propagator_gradient():
    """vp gradient through the drop-in propagator with the fused misfit equals the gradient
    with the reference's torch expressions of the same misfit."""
    from adfwi_amd.fwi import NormalizedWaveformL2, normalize_traces
    from adfwi_amd.model import AcousticModel
    from adfwi_amd.propagator import AcousticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    from oracle import misfit as om
    rng = np.random.default_rng(7)
    nz, nx, nt, dt = 40, 80, 300, 1e-3
    vp = (2200 + 600 * rng.random((nz, nx))).astype(np.float32)
    t = np.arange(nt) * dt - 0.08
    wav = (1 - 2 * (np.pi * 15 * t) ** 2) * np.exp(-(np.pi * 15 * t) ** 2)
    src = Source(nt, dt, 15.0)
    src.add_sources([10, 40, 70], [2, 2, 2], wav)
    rcv = Receiver(nt, dt)
    rcv.add_receivers(np.arange(nx), np.full(nx, 3), "pr")
    obs = normalize_traces(torch.tensor(rng.standard_normal((3, nt, nx)).astype(np.float32), device=DEV))
    grads = []
    for fused in (True, False):
        model = AcousticModel(0, 0, nx, nz, 10.0, 10.0, vp=vp, rho=np.full((nz, nx), 2000.0, np.float32),
                              vp_grad=True, nabc=10, free_surface=False, device=DEV)
        prop = AcousticPropagator(model, Survey(src, rcv), device=DEV)
        rec = prop.forward()["p"]
        loss = NormalizedWaveformL2(dt).forward(rec, obs) if fused else om.driver_misfit(rec, obs, dt, True)
        loss.backward()
        grads.append(model.vp.grad.cpu().numpy())
    assert rel_l2(grads[0], grads[1]) <= 1e-5, rel_l2(grads[0], grads[1])


@pytest.mark.parametrize("cfg", [
    dict(grad_mute=8, grad_smooth=0, marine_or_land="land"),
    dict(grad_mute=8, grad_smooth=5, marine_or_land="marine"),
    dict(grad_mute=6, grad_smooth=3, marine_or_land="Marine", mask=True),
    dict(grad_mute=0, grad_smooth=4, marine_or_land="onshore", norm_grad=False),
])
def test_grad_processor_on_device_matches_oracle(cfg):
    from adfwi_amd.fwi import GradProcessor
    from oracle import gradient_process as ogp
    rng = np.random.default_rng(11)
    nz, nx = 60, 90
    grad = rng.standard_normal((nz, nx)).astype(np.float32)
    forw = np.abs(rng.standard_normal((nz, nx))).astype(np.float32) * np.linspace(1, 0.01, nz)[:, None]
    cfg = dict(cfg)
    mask = rng.uniform(0.5, 1.0, (nz, nx)) if cfg.pop("mask", False) else None
    kw = dict(grad_mask=mask, norm_grad=cfg.pop("norm_grad", True), forw_illumination=True, **cfg)
    ref = ogp.GradProcessor(**kw).forward(nx, nz, 3500.0, grad.copy(), forw=forw)
    got_np = GradProcessor(device=DEV, **kw).forward(nx, nz, 3500.0, grad.copy(), forw=forw)
    assert isinstance(got_np, np.ndarray)
    assert np.allclose(got_np, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    got_t = GradProcessor(device=DEV, **kw).forward(nx, nz, 3500.0, torch.tensor(grad, device=DEV),
                                                    forw=torch.tensor(forw, device=DEV))
    assert got_t.dtype == torch.float32 and got_t.device.type == "cuda"
    assert rel_l2(got_t.cpu().numpy(), ref) <= 1e-6
