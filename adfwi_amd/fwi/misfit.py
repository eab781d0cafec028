"""Waveform L2 misfit and the drivers' per-trace normalisation on the GPU.

`Misfit_waveform_L2(dt).forward(obs, syn)` keeps the reference's class, signature and value
(ADFWI/fwi/misfit/L2.py:13-29: sum over traces of sqrt(sum_t (obs - syn)^2 dt)).
`Misfit_waveform_L2_squared(dt).forward(obs, syn)` is the least-squares objective
1/2 dt sum (obs - syn)^2 (no counterpart in the reference's misfit set): the reference's L2 takes
a square root per trace, whose gradient is NaN on a trace with zero residual -- a far-offset
trace that no arrival reaches within the record has exactly zero residual.
`NormalizedWaveformL2(dt).forward(rec, obs)` is what AcousticFWI computes per batch
(fwi/acoustic_fwi.py:148-153): the synthetic records normalised per trace by their maximum
absolute value, then that L2 against the normalised observed records -- in one kernel
(`adfwi_trace_l2`) that also writes the adjoint source through the normalisation, so the
records' gradient needs no further pass.  Both are torch.autograd Functions of the records.
"""
import torch

from .. import _native as nat


def normalize_traces(x: torch.Tensor) -> torch.Tensor:
    """x / max_t |x| per trace of (S, nt, R) records (fwi/acoustic_fwi.py:68-69, applied once
    to the observed data)."""
    return x / torch.max(torch.abs(x), dim=1, keepdim=True).values


NORMALIZE, SQUARED = 1, 2   # adfwi_trace_l2 flags (ADFWI_L2_NORMALIZE, ADFWI_L2_SQUARED)


def _trace_l2(a: torch.Tensor, b: torch.Tensor, dt: float, flags: int, want_a: bool, want_b: bool):
    if a.shape != b.shape or a.dim() != 3:
        raise ValueError(f"records must both be (shots, nt, receivers); got {tuple(a.shape)} and {tuple(b.shape)}")
    nat.require_gpu(a.device)
    lib = nat.load()
    a = a.detach().to(torch.float32).contiguous()
    b = b.detach().to(device=a.device, dtype=torch.float32).contiguous()
    S, nt, R = a.shape
    loss = torch.empty((S, R), device=a.device, dtype=torch.float32)
    ga = torch.empty_like(a) if want_a else None
    gb = torch.empty_like(b) if want_b else None
    nat.check(lib.adfwi_trace_l2(a.data_ptr(), b.data_ptr(), S, nt, R, float(dt), int(flags), loss.data_ptr(),
                                 None if ga is None else ga.data_ptr(), None if gb is None else gb.data_ptr(),
                                 nat.stream_handle(a.device)), "adfwi_trace_l2")
    return loss, ga, gb


class _TraceL2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, dt, flags):
        loss, ga, gb = _trace_l2(a, b, dt, flags, ctx.needs_input_grad[0], ctx.needs_input_grad[1])
        ctx.save_for_backward(*(t for t in (ga, gb) if t is not None))
        ctx.has = (ga is not None, gb is not None)
        return loss.sum()

    @staticmethod
    def backward(ctx, g):
        saved = list(ctx.saved_tensors)
        ga = saved.pop(0) * g if ctx.has[0] else None
        gb = saved.pop(0) * g if ctx.has[1] else None
        return ga, gb, None, None


class Misfit_waveform_L2:
    """Waveform difference L2 norm (ADFWI/fwi/misfit/L2.py:13-29), on the GPU."""

    def __init__(self, dt=1) -> None:
        self.dt = dt

    def forward(self, obs: torch.Tensor, syn: torch.Tensor) -> torch.Tensor:
        return _TraceL2.apply(obs, syn, float(self.dt), 0)


class Misfit_waveform_L2_squared:
    """Least-squares waveform misfit 1/2 dt sum_{s,t,r} (obs - syn)^2, on the GPU (the same kernel
    as Misfit_waveform_L2).  No counterpart in the reference: its gradient is defined on traces
    with zero residual, where the per-trace square root of Misfit_waveform_L2 gives NaN."""

    def __init__(self, dt=1) -> None:
        self.dt = dt

    def forward(self, obs: torch.Tensor, syn: torch.Tensor) -> torch.Tensor:
        return _TraceL2.apply(obs, syn, float(self.dt), SQUARED)


class NormalizedWaveformL2:
    """AcousticFWI's batch misfit (fwi/acoustic_fwi.py:148-153) fused: L2 of the per-trace
    normalised synthetic records against normalised observed records."""

    def __init__(self, dt=1) -> None:
        self.dt = dt

    def forward(self, rec: torch.Tensor, obs: torch.Tensor) -> torch.Tensor:
        return _TraceL2.apply(rec, obs, float(self.dt), NORMALIZE)
