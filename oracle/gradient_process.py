"""Oracle: the reference's gradient post-processing in numpy / scipy, float64 (test
infrastructure only).

Restates ADFWI/propagator/gradient_process.py:
* `gauss_kernel` / `smooth2d` (:14-49): a (2 span + 1)^2 Gaussian sampled at x, y in
  [-2 span, 2 span] with spacing 2 and covariance span^2 I, normalised to unit sum; the field
  and an all-ones field are convolved with it ('same', zero fill) and divided;
* `grad_taper` (:51-71): 'Marine' / 'Offshore' (case-sensitive) zero the first `tapersize`
  rows; otherwise the second half of a Hamming window of length 2 tapersize is written into
  the first `tapersize` *columns* of every row (the loop variable runs over rows), smoothed
  with span tapersize // 2, scaled to max 1 - thred, turned into 1 - taper and squared.
  scipy.signal.hamming (the reference's call, scipy 1.10) is scipy.signal.windows.hamming in
  current scipy;
* `GradProcessor.forward` (:85-134): taper, mask, illumination preconditioning by the
  smoothed forward wavefield (span 40, or half the smaller dimension below 41), smoothing of
  the gradient (below the muted rows for marine data) and scaling to vmax / max |grad|.
"""
import numpy as np
import scipy.signal as sps
from scipy.signal import windows


def gauss_kernel(span: int) -> np.ndarray:
    ax = np.linspace(-2.0 * span, 2.0 * span, 2 * span + 1)
    xx, yy = np.meshgrid(ax, ax)
    inv = 1.0 / float(span) ** 2                  # inverse of the diagonal covariance span^2 I
    f = np.exp(-0.5 * (inv * xx ** 2 + inv * yy ** 2))
    return f / f.sum()


def smooth2d(z: np.ndarray, span: int = 10) -> np.ndarray:
    f = gauss_kernel(span)
    num = sps.convolve2d(np.array(z, dtype=np.float64), f, mode="same")
    den = sps.convolve2d(np.ones(np.shape(z)), f, mode="same")
    return num / den


def grad_taper(nz: int, nx: int, tapersize: int = 20, thred: float = 0.05, marine_or_land: str = "Marine"):
    if marine_or_land in ("Marine", "Offshore"):
        taper = np.ones((nz, nx))
        taper[:tapersize, :] = 0.0
        return taper
    half = windows.hamming(2 * tapersize)[tapersize:]
    taper = np.zeros((nz, nx))
    taper[:, :tapersize] = half[None, :]
    taper = smooth2d(taper, span=tapersize // 2)
    taper = taper / taper.max() * (1.0 - thred)
    taper = 1.0 - taper
    return taper * taper


class GradProcessor:
    def __init__(self, grad_mute=0, grad_smooth=0, grad_mask=None, norm_grad=True, forw_illumination=True,
                 marine_or_land="land"):
        self.grad_mute, self.grad_smooth, self.grad_mask = grad_mute, grad_smooth, grad_mask
        self.norm_grad, self.forw_illumination, self.marine_or_land = norm_grad, forw_illumination, marine_or_land

    def forward(self, nx, nz, vmax, grad, forw=None):
        kind = self.marine_or_land.lower()
        if kind in ("marine", "offshore"):
            thred = 0.0
        elif kind in ("land", "onshore"):
            thred = 0.001
        else:
            raise ValueError(f"not supported modeling marine_or_land: {self.marine_or_land}")
        grad = np.array(grad, dtype=np.float64)
        if self.grad_mute > 0:
            grad = grad * grad_taper(nz, nx, self.grad_mute, thred, self.marine_or_land)
        if self.grad_mask is not None:
            grad = grad * self.grad_mask
        span = 40 if min(nz, nx) > 40 else int(min(nz, nx) / 2)
        if self.forw_illumination and forw is not None:
            pre = smooth2d(forw, span)
            pre = pre / np.max(pre + 1e-5)
            pre[pre < 1e-4] = 1e-4
            grad = grad / pre ** 2
        if self.grad_smooth > 0:
            if self.marine_or_land in ("marine", "offshore"):
                grad[self.grad_mute:, :] = smooth2d(grad[self.grad_mute:, :], span=self.grad_smooth)
            else:
                grad = smooth2d(grad, span=self.grad_smooth)
        if self.norm_grad:
            grad = vmax * grad / np.abs(grad).max()
        return grad
