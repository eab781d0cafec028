"""The caller contract of the reference's FWI drivers, on the GPU against the oracle on the CPU.

* AcousticFWI's batch loop (ADFWI/fwi/acoustic_fwi.py:134-177): shots in `shot_index` batches,
  the synthetic traces normalised per trace, the waveform L2 misfit against normalised observed
  data, one backward per batch accumulating into vp.grad, the forward wavefield summed over the
  batches, then the gradient processor (taper, illumination, smoothing, scaling).
* DIP (ADFWI/dip/dip_acoustic_fwi.py:126-162): vp is the non-leaf output of a network that
  model.forward() regenerates; every batch registers a hook on it that replaces its gradient
  with the processed one before it flows into the network's weights.

The drivers' code is restated here once and run twice: with adfwi_amd's propagator, misfit and
gradient processor on the GPU, and with the oracle's forward_kernel, misfit and gradient
processor on the CPU.  Bars: losses rel 1e-5, processed gradients and network-weight gradients
rel-L2 1e-4.
"""
import math

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

NZ, NX, NABC, NT, DT, DX, F0 = 40, 90, 10, 500, 1e-3, 10.0, 15.0
SRC_X = np.array([8, 20, 33, 45, 58, 70, 82])
BATCH = 3


def _ricker():
    t = np.arange(NT) * DT - 1.2 / F0
    tau = (np.pi * F0) ** 2
    return ((1 - 2 * tau * t ** 2) * np.exp(-tau * t ** 2)).astype(np.float32)


def _smooth(rng, lo, hi, passes=8):
    a = rng.standard_normal((NZ, NX))
    for _ in range(passes):
        a = 0.25 * (np.roll(a, 1, 0) + np.roll(a, -1, 0) + np.roll(a, 1, 1) + np.roll(a, -1, 1))
    a = (a - a.min()) / (a.max() - a.min())
    return (lo + (hi - lo) * a).astype(np.float32)


def _models():
    rng = np.random.default_rng(17)
    vp_true = _smooth(rng, 1800.0, 3200.0)
    vp_init = _smooth(rng, 2200.0, 2600.0, passes=30)
    return vp_true, vp_init


class _Runner:
    """One side of the comparison: `forward(vp, rho, shots)` -> record dict, `misfit`, and the
    gradient processor, either the product's (GPU) or the oracle's (CPU)."""

    def __init__(self, product, vmax):
        from adfwi_amd.propagator.boundary_condition import bc_pml
        self.product = product
        self.dev = DEV if product else "cpu"
        self.damp = torch.tensor(bc_pml(NX, NZ, DX, DX, NABC, vmax, False), dtype=torch.float32)
        if product:
            from adfwi_amd.fwi import GradProcessor, Misfit_waveform_L2
            from adfwi_amd.propagator import acoustic_kernels as mod
        else:
            from oracle import acoustic as mod
            from oracle.gradient_process import GradProcessor
            from oracle.misfit import waveform_l2

            class Misfit_waveform_L2:
                def __init__(self, dt):
                    self.dt = dt

                def forward(self, obs, syn):
                    return waveform_l2(obs, syn, self.dt)
        self.mod = mod
        self.loss_fn = Misfit_waveform_L2(DT)
        self.gp = GradProcessor(grad_mute=4, grad_smooth=3, marine_or_land="land")

    def forward(self, vp, rho, shots):
        d = self.dev
        S = len(shots)
        return self.mod.forward_kernel(NX, NZ, DX, DX, NT, DT, NABC, False,
                                       torch.tensor(SRC_X[shots], device=d), torch.full((S,), 2, device=d), S,
                                       torch.tensor(np.tile(_ricker(), (S, 1)), device=d),
                                       torch.arange(NX, device=d), torch.full((NX,), 2, device=d), NX,
                                       self.damp.to(d), vp, rho, device=d)

    def observed(self, vp_true):
        v = torch.tensor(vp_true, device=self.dev)
        with torch.no_grad():
            p = self.forward(v, 310 * v ** 0.25, np.arange(len(SRC_X)))["p"]
        return p / torch.max(torch.abs(p), dim=1, keepdim=True).values     # acoustic_fwi.py:66-69


def _batches(n):
    nb = math.ceil(n / BATCH)
    for b in range(nb):
        yield np.arange(b * BATCH, n if b == nb - 1 else (b + 1) * BATCH)


def _acoustic_fwi_iteration(run, obs, vp_init):
    """acoustic_fwi.py:134-177 for one iteration, waveform_normalize=True."""
    vp = torch.tensor(vp_init, device=run.dev, requires_grad=True)
    loss_batch, forw = 0.0, None
    for shots in _batches(len(SRC_X)):
        rho = (310 * vp.detach() ** 0.25)           # the model's Gardner update from detached vp
        rec = run.forward(vp, rho, shots)
        fw = rec["forward_wavefield_p"].cpu().detach().numpy()
        forw = fw if forw is None else forw + fw
        syn_p = rec["p"]
        syn_p = syn_p / (torch.max(torch.abs(syn_p), axis=1, keepdim=True).values)
        data_loss = run.loss_fn.forward(syn_p, obs[torch.as_tensor(shots, device=obs.device)])
        loss_batch += data_loss.item()
        data_loss.backward()
    grads = vp.grad.cpu().detach().numpy()
    vmax = np.max(vp.cpu().detach().numpy())
    grads = run.gp.forward(nz=NZ, nx=NX, vmax=vmax, grad=grads, forw=forw)
    return loss_batch, np.asarray(grads.cpu() if torch.is_tensor(grads) else grads, dtype=np.float64), forw


def test_acoustic_fwi_batch_loop_matches_oracle():
    vp_true, vp_init = _models()
    vmax = float(vp_init.max())
    out = {}
    for product in (True, False):
        run = _Runner(product, vmax)
        out[product] = _acoustic_fwi_iteration(run, run.observed(vp_true), vp_init)
    (l_h, g_h, f_h), (l_o, g_o, f_o) = out[True], out[False]
    assert abs(l_h - l_o) <= 1e-5 * abs(l_o), (l_h, l_o)
    assert rel_l2(f_h, f_o) <= 1e-5, rel_l2(f_h, f_o)
    assert np.isfinite(g_h).all() and np.abs(g_h).max() > 0
    assert rel_l2(g_h, g_o) <= 1e-4, rel_l2(g_h, g_o)


class _Generator(torch.nn.Module):
    """A small deep-image-prior network in the shape of ADFWI/dip/model/CNN.py: a fixed latent
    vector through a linear layer, bilinear upsampling and convolutions, tanh-scaled to
    [vmin, vmax] (km/s, times 1000) and cropped to the model grid."""

    def __init__(self, vmin=2.0, vmax=3.2):
        super().__init__()
        torch.manual_seed(1234)
        self.vmin, self.vmax = vmin, vmax
        self.h, self.w = math.ceil(NZ / 4), math.ceil(NX / 4)
        self.fc = torch.nn.Linear(16, self.h * self.w * 4, bias=False)
        self.conv1 = torch.nn.Conv2d(4, 8, 3, padding=1, bias=False)
        self.conv2 = torch.nn.Conv2d(8, 1, 3, padding=1, bias=False)
        self.z = torch.rand(16)

    def forward(self):
        out = torch.nn.functional.leaky_relu(self.fc(self.z.to(self.fc.weight.device)), 0.1)
        out = out.view(1, 4, self.h, self.w)
        out = torch.nn.functional.interpolate(out, scale_factor=4, mode="bilinear", align_corners=True)
        out = torch.nn.functional.leaky_relu(self.conv1(out), 0.1)
        out = self.conv2(out).squeeze()
        out = ((self.vmax - self.vmin) * torch.tanh(out) + (self.vmax + self.vmin)) / 2 * 1000
        return out[:NZ, :NX]


def _dip_iteration(run, obs, net):
    """dip_acoustic_fwi.py:126-162 for one iteration: vp regenerated per batch (model.forward()),
    a gradient hook on the non-leaf vp, normalised traces, L2 misfit, backward per batch."""
    net.zero_grad()
    forw, loss_batch = None, 0.0
    for shots in _batches(len(SRC_X)):
        vp = net()                                                  # DIP model.forward(): non-leaf vp
        rho = 310 * vp.detach() ** 0.25
        rec = run.forward(vp, rho, shots)
        fw = rec["forward_wavefield_p"].cpu().detach().numpy()
        forw = fw if forw is None else forw + fw
        snapshot = forw.copy()

        def grad_post_process(grads, snapshot=snapshot, vp=vp):
            g = grads.cpu().detach().numpy()
            with torch.no_grad():
                vmax = np.max(vp.cpu().detach().numpy())
                g = run.gp.forward(nz=NZ, nx=NX, vmax=vmax, grad=g, forw=snapshot)
                g = torch.as_tensor(np.asarray(g.cpu() if torch.is_tensor(g) else g), dtype=torch.float32)
            return g.to(grads.device)
        assert not vp.is_leaf and vp.requires_grad
        vp.register_hook(grad_post_process)
        syn_p = rec["p"]
        syn_p = syn_p / (torch.max(torch.abs(syn_p), axis=1, keepdim=True).values)
        data_loss = run.loss_fn.forward(syn_p, obs[torch.as_tensor(shots, device=obs.device)])
        loss_batch += data_loss.item()
        data_loss.backward()
    return loss_batch, {n: p.grad.detach().cpu().double().numpy() for n, p in net.named_parameters()}


def test_dip_non_leaf_vp_with_gradient_hook_matches_oracle():
    vp_true, _ = _models()
    out = {}
    for product in (True, False):
        net = _Generator()
        if product:
            net = net.to(DEV)
        with torch.no_grad():
            vmax = float(net().max())
        run = _Runner(product, vmax)
        out[product] = _dip_iteration(run, run.observed(vp_true), net)
    (l_h, g_h), (l_o, g_o) = out[True], out[False]
    assert abs(l_h - l_o) <= 1e-5 * abs(l_o), (l_h, l_o)
    for n in g_o:
        assert np.isfinite(g_h[n]).all() and np.abs(g_h[n]).max() > 0, n
        assert rel_l2(g_h[n], g_o[n]) <= 1e-4, (n, rel_l2(g_h[n], g_o[n]))
