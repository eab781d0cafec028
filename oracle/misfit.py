"""Oracle: the FWI drivers' per-trace normalisation and the waveform L2 misfit in eager
PyTorch (test infrastructure only).

Restates ADFWI/fwi/acoustic_fwi.py:68-69 and :149-150 (every trace divided by its maximum
absolute value over time, axis 1 of the (S, nt, R) records) and ADFWI/fwi/misfit/L2.py:26-29
(rsd = obs - syn; sum over traces of sqrt(sum_t rsd * rsd * dt)).  The driver passes its
arguments in the order (synthetic, observed), acoustic_fwi.py:152-153, so rsd = syn - obs
there; the misfit value is symmetric.  Gradients come from torch autograd, as in the reference.
"""
import torch


def normalize_traces(x: torch.Tensor) -> torch.Tensor:
    """acoustic_fwi.py:68-69 / :149-150: x / max_t |x| per (shot, receiver) trace."""
    peak = torch.max(torch.abs(x), dim=1, keepdim=True).values
    return x / peak


def waveform_l2(obs: torch.Tensor, syn: torch.Tensor, dt: float) -> torch.Tensor:
    """fwi/misfit/L2.py:26-29."""
    rsd = obs - syn
    per_trace = torch.sum(rsd * rsd * dt, dim=1)
    return torch.sum(torch.sqrt(per_trace))


def waveform_l2_squared(obs: torch.Tensor, syn: torch.Tensor, dt: float) -> torch.Tensor:
    """The least-squares objective 1/2 sum (obs - syn)^2 dt (no counterpart in the reference; the
    product's Misfit_waveform_L2_squared).  Pinned by its closed form: on traces where
    waveform_l2's per-trace root r is positive, this equals 1/2 sum r^2."""
    rsd = obs - syn
    return 0.5 * torch.sum(rsd * rsd * dt)


def driver_misfit(rec: torch.Tensor, obs: torch.Tensor, dt: float, normalize: bool = True) -> torch.Tensor:
    """The misfit of one batch as AcousticFWI forms it (acoustic_fwi.py:148-153): the synthetic
    records normalised per trace when waveform_normalize, then L2 against the (already
    normalised) observed records, synthetic first."""
    syn = normalize_traces(rec) if normalize else rec
    return waveform_l2(syn, obs, dt)
