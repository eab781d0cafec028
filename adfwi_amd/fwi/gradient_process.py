"""Gradient post-processing on the GPU (ADFWI/propagator/gradient_process.py:14-134).

Same constructor, `forward(nx, nz, vmax, grad, forw)` and results as the reference's
GradProcessor, computed in float64 on the device instead of numpy on the host:
* `smooth2d` (:29-49): the reference's Gaussian (2 span + 1 taps per axis at spacing 2,
  covariance span^2 I) is separable, so the 'same' zero-filled 2-D convolution is two 1-D
  convolutions; the field is divided by the same smoothing of an all-ones field;
* `grad_taper` (:51-71): 'Marine' / 'Offshore' (case-sensitive, as the reference tests them)
  zero the first `tapersize` rows; any other value takes the land branch, which writes the
  second half of a Hamming window into the first `tapersize` columns of every row, smooths
  it, scales it to 1 - thred, and returns (1 - taper)^2;
* `GradProcessor.forward` (:85-134): taper, mask, illumination preconditioning by the
  smoothed forward wavefield, gradient smoothing (below the muted rows for 'marine' /
  'offshore' in lower case, as the reference compares), scaling to vmax / max |grad|.
numpy inputs give numpy results (the reference driver's calling convention,
fwi/acoustic_fwi.py:171-175); torch inputs give float32 tensors on their device.
"""
import math
from typing import Optional

import numpy as np
import torch
import torch.nn.functional as F

from .. import _native as nat


def _gauss_taps(span: int, device) -> torch.Tensor:
    ax = torch.linspace(-2.0 * span, 2.0 * span, 2 * span + 1, dtype=torch.float64, device=device)
    g = torch.exp(-0.5 * ax * ax / float(span) ** 2)
    return g / g.sum()


def _conv_same(z: torch.Tensor, taps: torch.Tensor) -> torch.Tensor:
    """'same' zero-filled convolution of a (nz, nx) field with outer(taps, taps)."""
    k = taps.numel()
    p = k // 2
    x = z[None, None]
    x = F.conv2d(F.pad(x, (p, p, 0, 0)), taps.flip(0).view(1, 1, 1, k))
    x = F.conv2d(F.pad(x, (0, 0, p, p)), taps.flip(0).view(1, 1, k, 1))
    return x[0, 0]


def smooth2d(z: torch.Tensor, span: int = 10) -> torch.Tensor:
    """gradient_process.py:29-49 on the device (float64)."""
    z = z.to(torch.float64)
    taps = _gauss_taps(span, z.device)
    return _conv_same(z, taps) / _conv_same(torch.ones_like(z), taps)


def _hamming(n: int, device) -> torch.Tensor:
    """Symmetric Hamming window of length n: 0.54 - 0.46 cos(2 pi k / (n - 1))."""
    k = torch.arange(n, dtype=torch.float64, device=device)
    return 0.54 - 0.46 * torch.cos(2.0 * math.pi * k / (n - 1))


def grad_taper(nz: int, nx: int, tapersize: int = 20, thred: float = 0.05, marine_or_land: str = "Marine",
               device="cuda:0") -> torch.Tensor:
    """gradient_process.py:51-71 on the device (float64)."""
    if marine_or_land in ("Marine", "Offshore"):
        taper = torch.ones((nz, nx), dtype=torch.float64, device=device)
        taper[:tapersize, :] = 0.0
        return taper
    half = _hamming(2 * tapersize, device)[tapersize:]
    taper = torch.zeros((nz, nx), dtype=torch.float64, device=device)
    taper[:, :tapersize] = half[None, :]
    taper = smooth2d(taper, span=tapersize // 2)
    taper = taper / taper.max() * (1.0 - thred)
    taper = 1.0 - taper
    return taper * taper


class GradProcessor:
    """gradient_process.py:74-134 on the GPU."""

    def __init__(self, grad_mute=0, grad_smooth=0, grad_mask=None, norm_grad=True, forw_illumination=True,
                 marine_or_land="land", device="cuda:0"):
        self.grad_mute = grad_mute
        self.grad_smooth = grad_smooth
        self.grad_mask = grad_mask
        self.marine_or_land = marine_or_land
        self.norm_grad = norm_grad
        self.forw_illumination = forw_illumination
        self.device = device

    def forward(self, nx, nz, vmax, grad, forw=None):
        as_numpy = isinstance(grad, np.ndarray)
        dev = torch.device(self.device) if as_numpy else grad.device
        nat.require_gpu(dev)
        kind = self.marine_or_land.lower()
        if kind in ("marine", "offshore"):
            thred = 0.0
        elif kind in ("land", "onshore"):
            thred = 0.001
        else:
            raise ValueError("not supported modeling marine_or_land: %s" % (self.marine_or_land))
        g = torch.as_tensor(grad, device=dev).to(torch.float64)
        if self.grad_mute > 0:
            g = g * grad_taper(nz, nx, tapersize=self.grad_mute, thred=thred,
                               marine_or_land=self.marine_or_land, device=dev)
        if self.grad_mask is not None:
            mask = torch.as_tensor(self.grad_mask, device=dev).to(torch.float64)
            if tuple(mask.shape) != tuple(g.shape):
                raise ValueError("Wrong size of grad mask: the size of the mask should be identical to the size "
                                 "of vp model")
            g = g * mask
        span = 40 if min(nz, nx) > 40 else int(min(nz, nx) / 2)
        if self.forw_illumination and forw is not None:
            pre = smooth2d(torch.as_tensor(forw, device=dev), span)
            pre = pre / torch.max(pre + 1e-5)
            pre = torch.clamp(pre, min=1e-4)
            g = g / pre ** 2
        if self.grad_smooth > 0:
            if self.marine_or_land in ("marine", "offshore"):
                g = g.clone()
                g[self.grad_mute:, :] = smooth2d(g[self.grad_mute:, :], span=self.grad_smooth)
            else:
                g = smooth2d(g, span=self.grad_smooth)
        if self.norm_grad:
            g = vmax * g / g.abs().max()
        if as_numpy:
            return g.cpu().numpy()
        return g.to(torch.float32)
