"""Hand-written elastic adjoint: segment recomputation from the sqrt(nt) checkpoints into a
per-segment trace of forward states, then the reverse steps of libadfwi_amd.so.

Replaces the autograd replay of the reference's checkpointed chunks
(ADFWI/propagator/elastic_kernels.py:975-1008, step_forward_PML_4order :222-423).
"""
import ctypes
from typing import List, Optional, Sequence

import torch

from .. import _native as nat
from .acoustic_kernels import TIMELINE

NCOEF = 8


def run_backward(geom, coef: torch.Tensor, src: torch.Tensor, pml_x: Optional[torch.Tensor],
                 pml_z: Optional[torch.Tensor], scalars, ckpt: Optional[torch.Tensor], k: int,
                 g_rec: Sequence[Optional[torch.Tensor]], needs: Sequence[bool], damp: Optional[torch.Tensor] = None,
                 traces: Optional[dict] = None):
    """Adjoint of all nt steps.

    ckpt: the forward's checkpoints (slot j = state after step (j+1)k-1), g_rec: five (S, nt, R)
    record adjoints (txx, tzz, txz, vx, vz; None = zero).  traces: {segment: per-step states}
    kept by run_forward(resident=...); those segments are not recomputed, and each is released
    once the segment before it no longer needs its last state.  Returns (list of 8 (Z, pitch)
    gradients of the padded coefficient planes C11, C13, C33, C55, bx, bz, C15, C35 -- None
    where not needed -- and the (S, nt) wavelet gradient or None).
    """
    lib = nat.load()
    dev = geom.device
    S, Z, nt, pitch = geom.S, geom.Z, geom.nt, geom.pitch
    L = geom.layout
    nseg = -(-nt // k)
    traces = {} if traces is None else traces
    first_resident = min(traces) if traces else nseg
    if first_resident > 1 and (ckpt is None or ckpt.shape[0] < first_resident - 1):
        raise RuntimeError("the forward pass kept no checkpoints for the adjoint")
    lam = geom.state(2)
    trace = geom.state(min(k, nt)) if len(traces) < nseg else None
    zero = None
    gacc = torch.zeros(int(L.accum_floats), device=dev, dtype=torch.float32)
    gsrc = torch.zeros((S, nt), device=dev, dtype=torch.float32)
    g = [None if x is None else x.detach().to(torch.float32).contiguous() for x in g_rec]
    d = geom.desc(coef, pml_x, pml_z, src, scalars, damp)
    stream = nat.stream_handle(dev)
    cur = lam[0]
    for seg in reversed(range(nseg)):
        t0 = seg * k
        n = min(nt, t0 + k) - t0
        if seg == 0:
            if zero is None:
                zero = geom.state(1)[0]
            fwd0 = zero
        elif seg - 1 in traces:
            fwd0 = traces[seg - 1][-1]     # the last state of the resident segment before
        else:
            fwd0 = ckpt[seg - 1]
        seg_trace = traces.get(seg)
        if seg_trace is None:
            seg_trace = trace
            fa = nat.ElasticFwdArgs()
            fa.state_in = fwd0.data_ptr()
            fa.trace = trace.data_ptr()
            fa.t0, fa.nsteps = t0, n
            TIMELINE.call("recompute", geom.kernel_names[0], n, dev, lambda: nat.check(
                lib.adfwi_elastic_forward(ctypes.byref(d), ctypes.byref(fa), stream),
                "adfwi_elastic_forward (recompute)"))
        aa = nat.ElasticAdjArgs()
        aa.lam_in = cur.data_ptr()
        aa.work[0], aa.work[1] = lam[0].data_ptr(), lam[1].data_ptr()
        aa.t_hi, aa.nsteps = t0 + n, n
        aa.fwd0, aa.trace = fwd0.data_ptr(), seg_trace.data_ptr()
        for i in range(5):
            aa.grad_rec[i] = None if g[i] is None else g[i].data_ptr()
        aa.grad_coef = gacc.data_ptr()
        aa.grad_src = gsrc.data_ptr()
        TIMELINE.call("adjoint", geom.kernel_names[1], n, dev, lambda: nat.check(
            lib.adfwi_elastic_adjoint(ctypes.byref(d), ctypes.byref(aa), stream), "adfwi_elastic_adjoint"))
        cur = lam[0] if aa.lam_out == lam[0].data_ptr() else lam[1]
        traces.pop(seg, None)              # done; segment seg + 1, which read its last state, too
        seg_trace = fwd0 = None
    traces.clear()
    planes = gacc.view(-1, NCOEF, Z, pitch).sum(0)
    gcoef: List[Optional[torch.Tensor]] = [planes[i] if (i < len(needs) and needs[i]) else None
                                           for i in range(NCOEF)]
    return gcoef, (gsrc if len(needs) > 8 and needs[8] else None)
