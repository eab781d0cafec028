"""Acoustic forward modelling replayed from HIP graphs.

A forward run without checkpoints (modelling requests, the observed-data forward) is the
same short launch sequence repeated: per pair of steps one two-step kernel and one receiver
gather.  `GraphedForward` captures the launches of one window of `segment` steps once, with
the library's device-side time base (`adfwi_acoustic_fwd_args.t_base`) as the only thing that
changes between windows, and ends the window with a node that advances the time base; the
run is then `nt // segment` replays of that graph plus one graph for the last, shorter
window.  The same launches run in the same order as in `run_forward`, so records and energy
maps are bit-identical to it.

A context owns static buffers (coefficients, wavelets, source positions, state, records,
energy) and its graphs; `graphed_forward` keeps a few contexts per host thread, keyed by the
geometry's content (grid, nt, shot count, receivers, tiling), so that later runs of the same
shape only copy their inputs in and replay.  A shape is captured on its second run (the first
runs directly), so one-off shapes pay no capture.  Graphs record no memset: buffers are
cleared outside the graphs, and the time-base update is an elementwise kernel.
"""
import copy
import ctypes
import threading
from collections import OrderedDict
from typing import Optional

import torch

from .. import _native as nat

_LOCAL = threading.local()
MAX_CONTEXTS = 4                   # per host thread
MAX_CONTEXT_BYTES = 256 << 20      # larger contexts are not kept (built, run once and dropped)


class GraphedForward:
    def __init__(self, geom, segment: int, energy_u_from: Optional[int]):
        if segment < 4 or segment % 4:
            raise ValueError("the graph window must be a positive multiple of 4 steps")
        dev = geom.device
        # a private copy of the device index tables: replays write the source positions of
        # each run into it, never into a geometry that a propagator still holds
        geom = copy.copy(geom)
        geom.src_z, geom.src_x = geom.src_z.clone(), geom.src_x.clone()
        geom.rcv_z, geom.rcv_x = geom.rcv_z.clone(), geom.rcv_x.clone()
        geom.rcv_map = None if geom.rcv_map is None else geom.rcv_map.clone()
        self.geom, self.K = geom, int(segment)
        S, Z, X, nt, R, pitch = geom.S, geom.Z, geom.X, geom.nt, geom.R, geom.pitch
        L = geom.layout
        self.nfull, self.rem = nt // self.K, nt % self.K
        self.coef = torch.zeros((5, Z, pitch), device=dev, dtype=torch.float32)
        self.src = torch.zeros((S, nt), device=dev, dtype=torch.float32)
        self.work = geom.state(2)
        self.rec = [torch.zeros((S, nt, R), device=dev, dtype=torch.float32) for _ in range(3)] if R else None
        self.energy = (torch.zeros((L.ngroups, 2, Z, pitch), device=dev, dtype=torch.float32)
                       if energy_u_from is not None else None)
        self.eu_from = nt if energy_u_from is None else int(energy_u_from)
        self.tb = torch.zeros(1, device=dev, dtype=torch.int32)
        self.nbytes = sum(t.numel() * 4 for t in [self.coef, self.src, self.work]
                          + (self.rec or []) + ([self.energy] if self.energy is not None else []))
        lib = nat.load()
        self._desc = geom.desc(self.coef, self.src)
        self._graphs = []
        final = self.work[0]
        side = torch.cuda.Stream(device=dev)                 # default priority
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for n, advance in ((self.K, True), (self.rem, False)):
                if n == 0 or (advance and self.nfull == 0):
                    continue
                a = self._args(n)
                g = torch.cuda.CUDAGraph()
                g.capture_begin(capture_error_mode="thread_local")
                try:
                    nat.check(lib.adfwi_acoustic_forward(ctypes.byref(self._desc), ctypes.byref(a),
                                                         nat.stream_handle(dev)), "adfwi_acoustic_forward (graph)")
                    if advance:
                        self.tb.add_(self.K)
                finally:
                    g.capture_end()
                out = a.state_out
                if advance and out != self.work[0].data_ptr():
                    raise RuntimeError("a graph window must leave the state in work[0]")
                if not advance:
                    final = self.work[0] if out == self.work[0].data_ptr() else self.work[1]
                self._graphs.append((g, advance))
        torch.cuda.current_stream(dev).wait_stream(side)
        self.final = final

    def _args(self, n: int) -> nat.AcousticFwdArgs:
        g = self.geom
        a = nat.AcousticFwdArgs()
        a.state_in = self.work[0].data_ptr()
        a.work[0], a.work[1] = self.work[0].data_ptr(), self.work[1].data_ptr()
        a.ckpt, a.ckpt_every = None, 0
        a.t0, a.nsteps = 0, int(n)
        a.energy_u_from = self.eu_from
        if self.rec is not None:
            a.rec_p, a.rec_u, a.rec_w = (r.data_ptr() for r in self.rec)
        a.energy = None if self.energy is None else self.energy.data_ptr()
        a.rcv_map = None if g.rcv_map is None else g.rcv_map.data_ptr()
        a.fuse_steps = g.fuse_steps
        a.t_base = self.tb.data_ptr()
        return a

    def __call__(self, coef: torch.Tensor, src: torch.Tensor, geom=None) -> dict:
        """Same result dictionary as run_forward(record=True, energy_u_from=..., ckpt_every=0);
        `geom` (same content key) supplies the source positions of this run."""
        self.coef.copy_(coef)
        self.src.copy_(src)
        if geom is not None:
            self.geom.src_z.copy_(geom.src_z)
            self.geom.src_x.copy_(geom.src_x)
        self.work[0].zero_()
        self.tb.zero_()
        if self.energy is not None:
            self.energy.zero_()
        for g, advance in self._graphs:
            for _ in range(self.nfull if advance else 1):
                g.replay()
        out = {"rec": None if self.rec is None else [r.clone() for r in self.rec], "ckpt": None, "energy": None}
        if self.energy is not None:
            X = self.geom.X
            e = self.energy.sum(0)[:, :, :X]
            e_w = self.final[2, :, :, :X].square().sum(0)
            out["energy"] = torch.stack([e[0], e[1], e_w])
        return out


def graphed_forward(geom, coef: torch.Tensor, src: torch.Tensor, energy_u_from: Optional[int], segment: int,
                    capture_after: int = 1):
    """run_forward(record=True, energy_u_from, ckpt_every=0) through a cached GraphedForward of
    this thread.  A shape (geometry content, window, energy cut) runs directly its first
    `capture_after` times and is captured on the next run; contexts over MAX_CONTEXT_BYTES
    are never built."""
    from .acoustic_kernels import run_forward
    cache = getattr(_LOCAL, "cache", None)
    if cache is None:
        cache = _LOCAL.cache = OrderedDict()
        _LOCAL.seen = OrderedDict()
    key = (geom.content_key, int(segment), energy_u_from, geom.device.index)
    ctx = cache.get(key)
    if ctx is None:
        seen = _LOCAL.seen
        n = seen.pop(key, 0)
        seen[key] = n + 1
        while len(seen) > 64:
            seen.popitem(last=False)
        if n < capture_after:
            return run_forward(geom, coef, src, record=True, energy_u_from=energy_u_from, ckpt_every=0)
        ctx = GraphedForward(geom, segment, energy_u_from)
        if ctx.nbytes > MAX_CONTEXT_BYTES:
            del ctx
            return run_forward(geom, coef, src, record=True, energy_u_from=energy_u_from, ckpt_every=0)
        cache[key] = ctx
        while len(cache) > MAX_CONTEXTS:
            _, old = cache.popitem(last=False)
            torch.cuda.current_stream(geom.device).synchronize()   # its graphs have finished
            del old
    else:
        cache.move_to_end(key)
    return ctx(coef, src, geom)


def clear_cache() -> None:
    """Drops this thread's graphed-forward contexts (after the stream has finished them)."""
    cache = getattr(_LOCAL, "cache", None)
    if cache:
        torch.cuda.synchronize()
        cache.clear()
    if getattr(_LOCAL, "seen", None) is not None:
        _LOCAL.seen.clear()
