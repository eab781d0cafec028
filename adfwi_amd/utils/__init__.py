"""Input helpers of the propagator path: numpy<->tensor conversion and source wavelets.

API of ADFWI/utils/utils.py:8-60 (numpy2tensor & co.) and ADFWI/utils/wavelets.py:4-36.
"""
import numpy as np
import torch


def numpy2tensor(a, dtype=torch.float32):
    """numpy/list -> tensor of `dtype`; tensors pass through unchanged (utils.py:8-16)."""
    if torch.is_tensor(a):
        return a
    return torch.tensor(a, requires_grad=False, dtype=dtype)


def tensor2numpy(a):
    return a.detach().cpu().numpy() if torch.is_tensor(a) else a


def gpu2cpu(a):
    return a.detach().cpu().numpy() if torch.is_tensor(a) else a


def list2numpy(a):
    return np.array(a)


def numpy2list(a):
    return list(np.asarray(a).tolist())


def wavelet(nt, dt, f0, amp0=1, t0=None, type="Ricker"):
    """Source time function (wavelets.py:4-36): returns (t, w).

    Ricker: w = amp0 (1 - 2 tau (t-t0)^2) exp(-tau (t-t0)^2), tau = (pi f0)^2,
    t0 = 1.2/f0 by default; 'Gaussian' is the Ricker integrated twice by
    cumulative sums; 'Ramp' is amp0/2 (1 + tanh(t/t0)).
    """
    t = np.arange(nt) * dt + 0.0
    t0 = t0 if t0 is not None else 1.2 / f0
    kind = type.lower()
    if kind in ("ricker", "gaussian"):
        tau = (np.pi * f0) ** 2
        w = amp0 * (1 - 2 * tau * (t - t0) ** 2) * np.exp(-tau * (t - t0) ** 2)
        if kind == "gaussian":
            w = np.cumsum(np.cumsum(w))
    elif kind == "ramp":
        w = amp0 * 0.5 * (1.0 + np.tanh(t / t0))
    else:
        raise ValueError("Support source types: Rikcer, Guassian, Ramp. \n\n"
                         f"Unknown source type: {type}")
    return t, w
