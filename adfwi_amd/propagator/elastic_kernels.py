"""Elastic propagation on the MI355X: coefficient set-up in torch, time stepping in HIP.

Drop-in for ADFWI/propagator/elastic_kernels.py: ``pad_torchSingle`` (:175-216),
``DiffCoef`` (:20-58) and ``forward_kernel`` (:917-1037, 4th-order PML branch) keep
their signatures and results.  The per-step op sequence of step_forward_PML_4order
(:222-423) runs as one fused HIP kernel per step; its autograd replay is replaced by
``ElasticWave``, whose backward is the hand-written discrete adjoint with sqrt(nt)
wavefield checkpoints and segment recomputation.
"""
import ctypes
import math
import os
from typing import Dict, List, Optional

import numpy as np
import torch

from .. import _native as nat
from .acoustic_kernels import (PITCH_ALIGN, TIMELINE, ResidentSet, checkpoint_interval, replicate_pad,
                               resident_budget_bytes)

NAMES = ("txx", "tzz", "txz", "vx", "vz")
NFIELD = 10


def DiffCoef(order: int, grid_scheme: str) -> torch.Tensor:
    """Staggered-grid FD coefficients solved in float32 (elastic_kernels.py:20-58)."""
    if grid_scheme == "r":
        b = torch.cat((torch.tensor([1 / 2]), torch.zeros(order - 1)))
        expo = lambda i, j: (j + 1) ** (2 * (i + 1) - 1)
    elif grid_scheme == "s":
        b = torch.cat((torch.tensor([1.0]), torch.zeros(order - 1)))
        expo = lambda i, j: (2 * (j + 1) - 1) ** (2 * (i + 1) - 1)
    else:
        raise ValueError("Invalid grid scheme. Use 'r' for rational or 's' for symmetric.")
    a = torch.tensor([[float(expo(i, j)) for j in range(order)] for i in range(order)], dtype=torch.float32)
    return torch.linalg.inv(a) @ b


def pad_torchSingle(data: torch.Tensor, pml: int, fs_offset: int, free_surface: bool = True,
                    device: torch.device = torch.device("cpu")) -> torch.Tensor:
    """Edge-replicating pad of an array by its own shape, float32 like the reference
    (elastic_kernels.py:175-216): fs_offset rows on top (plus pml without a free surface),
    pml rows below and pml columns on both sides."""
    top = fs_offset if free_surface else fs_offset + pml
    return replicate_pad(data.to(torch.float32), pml, pml, top, pml)


def chunk_ends(nt: int, segments: int) -> List[int]:
    """Last step of every torch.chunk(.., segments) time chunk (elastic_kernels.py:396)."""
    size = -(-nt // max(1, int(segments)))
    return [min(nt, s + size) - 1 for s in range(0, nt, size)]


class ElasticGeometry:
    """Device-side description of one elastic problem: padded grid, sources (with their
    moment-tensor factors and, optionally, vertical-force weights), receivers, tiling."""

    def __init__(self, Z: int, X: int, nt: int, free_surface: bool, src_z, src_x, src_m: torch.Tensor,
                 rcv_z, rcv_x, device, aniso: bool = False, shots_per_wave: int = 0, tile_z: int = 0,
                 damping: bool = False, src_fz: Optional[torch.Tensor] = None, nn: int = 2, c2: float = 0.0,
                 resident_traces: Optional[int] = None):
        nat.require_gpu(device)
        self.device = torch.device(device)
        self.Z, self.X, self.nt = int(Z), int(X), int(nt)
        self.free_surface = bool(free_surface)
        self.aniso = bool(aniso)
        self.damping = bool(damping)          # multiplicative damping scheme (5 fields) instead of PML
        self.nfield = 5 if self.damping else NFIELD
        # stencil half-order fd_order // 2: 2 for the 4th-order schemes, 3..5 for the three-term
        # PML scheme (window [nn, Z-nn-1], free-surface rows nn-2 .. nn+1), c2 its third coefficient
        self.nn, self.c2 = int(nn), float(c2)
        # checkpoint segments (the last ones) whose per-step states the forward pass keeps for the
        # backward instead of recomputing them there: None = as many as the memory budget holds,
        # 0 = none; ADFWI_RESIDENT_IMAGES sets it when the argument is None
        env = os.environ.get("ADFWI_RESIDENT_IMAGES")
        self.resident_traces = int(env) if env is not None and resident_traces is None else resident_traces
        # (forward, adjoint) step kernels, as the timeline and profiles name them
        self.kernel_names = (("el_abl_fwd", "el_abl_adj") if self.damping else
                             ("el3_fwd_step", "el3_adj_step") if self.nn > 2 else ("el_fwd_step", "el_adj_step"))
        self.pitch = -(-self.X // PITCH_ALIGN) * PITCH_ALIGN
        src_z = np.asarray(src_z, dtype=np.int64).reshape(-1)
        src_x = np.asarray(src_x, dtype=np.int64).reshape(-1)
        rcv_z = np.asarray(rcv_z, dtype=np.int64).reshape(-1)
        rcv_x = np.asarray(rcv_x, dtype=np.int64).reshape(-1)
        self.S, self.R = len(src_z), len(rcv_z)
        if self.S < 1:
            raise ValueError("at least one shot is required")
        w = self.nn
        for name, zz, xx in (("source", src_z, src_x), ("receiver", rcv_z, rcv_x)):
            if len(zz) and (zz.min() < w or zz.max() > self.Z - w - 1 or xx.min() < w or xx.max() > self.X - w - 1):
                raise ValueError(f"{name} index outside the update window [{w}, {self.Z - w - 1}] x "
                                 f"[{w}, {self.X - w - 1}] of the padded grid")
        i32 = lambda a: torch.tensor(a.astype(np.int32), device=self.device)
        self.src_z, self.src_x = i32(src_z), i32(src_x)
        self.rcv_z, self.rcv_x = i32(rcv_z), i32(rcv_x)
        # receivers per row, then the receiver index of every point: with it the 4th-order PML
        # step kernels take the records and add the record adjoints themselves
        self.rcv_map = self._receiver_map(rcv_z, rcv_x)
        self.src_m = src_m.detach().to(device=self.device, dtype=torch.float32).contiguous()
        # vertical-force weight per shot (None: moment-tensor sources only, as in the reference)
        self.src_fz = (None if src_fz is None else
                       torch.as_tensor(src_fz).detach().to(device=self.device, dtype=torch.float32).reshape(-1)
                       .contiguous())
        if self.src_fz is not None and self.src_fz.numel() != self.S:
            raise ValueError(f"src_fz needs one weight per shot ({self.S}), got {self.src_fz.numel()}")
        self.shots_per_wave = int(shots_per_wave or os.environ.get("ADFWI_EL_SHOTS_PER_WAVE", "0"))
        self.tile_z = int(tile_z or os.environ.get("ADFWI_EL_TILE_Z", "0"))
        self.layout = nat.Layout()
        d = self.desc(None, None, None, None, (0.0,) * 7)
        nat.check(nat.load().adfwi_elastic_get_layout(ctypes.byref(d), ctypes.byref(self.layout)),
                  "adfwi_elastic_get_layout")

    def _receiver_map(self, rcv_z: np.ndarray, rcv_x: np.ndarray) -> Optional[torch.Tensor]:
        """[Z] receivers per row, then [Z][pitch] receiver index (-1 elsewhere); None without
        receivers, when receivers repeat a point (their records are then taken and their adjoints
        added by separate launches), or when ADFWI_EL_FUSED_RECORDS=0."""
        if self.R == 0 or os.environ.get("ADFWI_EL_FUSED_RECORDS", "1") == "0":
            return None
        pts = rcv_z * self.pitch + rcv_x
        if len(np.unique(pts)) != len(pts):
            return None
        m = np.full(self.Z + self.Z * self.pitch, -1, dtype=np.int32)
        m[:self.Z] = np.bincount(rcv_z, minlength=self.Z)[:self.Z]
        m[self.Z + pts] = np.arange(self.R, dtype=np.int32)
        return torch.tensor(m, device=self.device)

    def desc(self, coef, pml_x, pml_z, src, scalars, damp=None) -> nat.ElasticDesc:
        """pml_x, pml_z: PML factor tables (PML scheme); damp: [Z][pitch] taper (damping scheme)."""
        d = nat.ElasticDesc()
        d.abi_version, d.struct_size = nat.ABI_VERSION, ctypes.sizeof(nat.ElasticDesc)
        d.nshots, d.nz, d.nx, d.pitch = self.S, self.Z, self.X, self.pitch
        d.free_surface, d.nt, d.nrec, d.aniso = int(self.free_surface), self.nt, self.R, int(self.aniso)
        d.shots_per_wave, d.tile_z = self.shots_per_wave, self.tile_z
        d.dt, d.dt_dx, d.dt_dz, d.dx, d.dz, d.c0, d.c1 = scalars
        d.damping = int(self.damping)
        d.nn, d.c2 = self.nn, self.c2
        ptr = lambda t: None if t is None else t.data_ptr()
        d.coef, d.pml_x, d.pml_z, d.src = ptr(coef), ptr(pml_x), ptr(pml_z), ptr(src)
        d.damp = ptr(damp)
        d.src_m = self.src_m.data_ptr()
        d.src_fz = None if self.src_fz is None else self.src_fz.data_ptr()
        d.src_z, d.src_x = self.src_z.data_ptr(), self.src_x.data_ptr()
        d.rcv_z, d.rcv_x = self.rcv_z.data_ptr(), self.rcv_x.data_ptr()
        d.rcv_map = ptr(self.rcv_map)
        return d

    def state(self, n: int = 1) -> torch.Tensor:
        return torch.zeros((n, self.nfield, self.S, self.Z, self.pitch), device=self.device, dtype=torch.float32)

    def pack_plane(self, f: torch.Tensor) -> torch.Tensor:
        out = torch.zeros((self.Z, self.pitch), device=self.device, dtype=torch.float32)
        out[:f.shape[0], :f.shape[1]] = f
        return out

    def pack_coef(self, fields) -> torch.Tensor:
        """fields: C11, C13, C33, C55, bx, bz, C15, C35 padded arrays (each from its own shape)."""
        coef = torch.zeros((8, self.Z, self.pitch), device=self.device, dtype=torch.float32)
        for k, f in enumerate(fields):
            if f is not None:
                coef[k, :f.shape[0], :f.shape[1]] = f
        return coef


def pml_factors(bcx_p: torch.Tensor, bcz_p: torch.Tensor, dt: float, pitch: int):
    """1 -/+ dt/2 BC and the reciprocal, as elastic_kernels.py:313-322 forms them, per column
    (x) and per row (z); the reference's profiles are constant along the other axis."""
    bx = bcx_p[2, :]
    bz = bcz_p[:, 2]
    if not (torch.equal(bcx_p, bx.expand_as(bcx_p)) and torch.equal(bcz_p, bz[:, None].expand_as(bcz_p))):
        raise ValueError("the PML profiles must vary along one axis only (as bc_pml_xz builds them)")
    px = torch.zeros((3, pitch), device=bcx_p.device, dtype=torch.float32)
    pxd = 1 + 0.5 * dt * bx
    px[0, :bx.shape[0]] = 1 - 0.5 * dt * bx
    px[1, :bx.shape[0]] = pxd
    px[2, :bx.shape[0]] = 1.0 / pxd
    pzd = 1 + 0.5 * dt * bz
    pz = torch.stack([1 - 0.5 * dt * bz, pzd, 1.0 / pzd]).contiguous()
    return px, pz


def _energy(state: torch.Tensor, X: int, free_surface: bool, nn: int = 2):
    """sum over shots of each total squared (elastic_kernels.py:414-418), free-surface tzz row nn = 0.
    A 5-field state (damping scheme) holds the totals themselves."""
    out = []
    for k in range(5):
        tot = state[k] if state.shape[0] == 5 else state[2 * k] + state[2 * k + 1]
        if k == 1 and free_surface:
            tot = tot.clone()
            tot[:, nn, :] = 0.0
        out.append(tot[:, :, :X].square().sum(0))
    return out


def resident_trace_segments(geom: ElasticGeometry, k: int) -> int:
    """How many of the last checkpoint segments keep their per-step states from the forward pass
    (the backward then skips their recomputation): at most geom.resident_traces (0 = none; None =
    as many as resident_budget_bytes allows once the checkpoints, the backward's recompute trace
    and its three states are set aside)."""
    nseg = -(-geom.nt // k)
    cap = geom.resident_traces
    if cap == 0:
        return 0
    state_bytes = int(geom.layout.state_floats) * 4
    seg_bytes = k * state_bytes
    if cap is None:
        cap = resident_budget_bytes(geom.device, (nseg + 3) * state_bytes + seg_bytes) // seg_bytes
    return int(max(0, min(cap, nseg)))


def _state_at(ptr: int, work, bufs) -> torch.Tensor:
    """The state tensor a forward call left its last step in (a work buffer, a checkpoint slot
    or a trace slot)."""
    for w in work:
        if w.data_ptr() == ptr:
            return w
    for buf in bufs:
        step = buf[0].numel() * buf.element_size()
        off = ptr - buf.data_ptr()
        if 0 <= off < buf.shape[0] * step and off % step == 0:
            return buf[off // step]
    raise RuntimeError("forward state pointer outside the work, checkpoint and trace buffers")


def run_forward(geom: ElasticGeometry, coef, pml_x, pml_z, src, scalars, *, record=True, energy_ends=None,
                ckpt_every: int = 0, damp=None, resident: int = 0):
    """All nt steps from a zero state.  Returns dict(rec=[5 x (S, nt, R)] or None,
    energy=[5 x (Z, X)] summed over the states after the steps in energy_ends, ckpt,
    traces={segment: per-step states}).

    resident > 0 (with ckpt_every): the last `resident` checkpoint segments write the state after
    each of their steps to a trace buffer of their own, which run_backward uses instead of
    recomputing the segment; their checkpoints are not written."""
    lib = nat.load()
    dev = geom.device
    S, nt, R = geom.S, geom.nt, geom.R
    work = geom.state(2)
    k = int(ckpt_every)
    nseg = -(-nt // k) if k else 1
    resident = int(resident) if k else 0
    head = (nseg - resident) * k if resident else nt      # steps before the resident segments
    ckpt = None
    if k and head > 0:
        ckpt = geom.state(max(nseg - 1, 1) if not resident else max(nseg - resident, 1))
    # A trace slot receives every window point [nn, Z-nn-1] x [nn, X-nn-1] of every field at
    # every step (and the free-surface rows the step writes); only the border outside the window
    # and the pitch padding need clearing.  ADFWI_POISON_RESIDENT=1 fills the rest with NaN first
    # (tests: a point the adjoint read without the forward writing it would show in the gradient).
    poison = os.environ.get("ADFWI_POISON_RESIDENT") == "1"

    def trace_buffer(n):
        t = torch.empty((n, geom.nfield, S, geom.Z, geom.pitch), device=dev, dtype=torch.float32)
        if poison:
            t.fill_(float("nan"))
        w = geom.nn
        t[..., :w, :].zero_()
        t[..., geom.Z - w:, :].zero_()
        t[..., :w].zero_()
        t[..., geom.X - w:].zero_()
        return t

    traces = ResidentSet((seg, trace_buffer(min(nt, (seg + 1) * k) - seg * k)) for seg in range(nseg - resident, nseg))
    if resident:
        state_bytes = int(geom.layout.state_floats) * 4
        traces.hold(3 * state_bytes + (k * state_bytes if resident < nseg else 0))
    rec = [torch.zeros((S, nt, R), device=dev) for _ in range(5)] if record and R else None
    d = geom.desc(coef, pml_x, pml_z, src, scalars, damp)
    ends = sorted(set(energy_ends or [])) or [nt - 1]
    # calls end at the energy-map steps and at the edges of the resident segments
    stops = set(ends) | {head - 1} | {min(nt, (seg + 1) * k) - 1 for seg in traces}
    energy = None
    cur = work[0]
    t0 = 0
    stream = nat.stream_handle(dev)
    bufs = ([] if ckpt is None else [ckpt]) + list(traces.values())
    for te in sorted(s for s in stops if s >= 0):
        a = nat.ElasticFwdArgs()
        a.state_in = cur.data_ptr()
        a.work[0], a.work[1] = work[0].data_ptr(), work[1].data_ptr()
        if t0 >= head:            # inside a resident segment: every state goes to its trace
            seg = t0 // k
            a.trace = traces[seg][t0 - seg * k].data_ptr()
        else:
            a.ckpt = None if ckpt is None else ckpt.data_ptr()
            a.ckpt_every = k if ckpt is not None else 0
        a.t0, a.nsteps = t0, te + 1 - t0
        if rec is not None:
            for i in range(5):
                a.rec[i] = rec[i].data_ptr()
        TIMELINE.call("forward" if t0 < head else "forward+trace", geom.kernel_names[0], a.nsteps, dev,
                      lambda: nat.check(lib.adfwi_elastic_forward(ctypes.byref(d), ctypes.byref(a), stream),
                                        "adfwi_elastic_forward"))
        cur = _state_at(a.state_out, work, bufs)
        if energy_ends is not None and te in ends:
            e = _energy(cur, geom.X, geom.free_surface, geom.nn)
            energy = e if energy is None else [x + y for x, y in zip(energy, e)]
        t0 = te + 1
    return {"rec": rec, "energy": energy, "ckpt": ckpt, "final": cur, "traces": traces}


class ElasticWave(torch.autograd.Function):
    """Five records (txx, tzz, txz, vx, vz) and energy maps from padded coefficient fields.

    forward(c11, c13, c33, c55, bx, bz, c15, c35, wavelet, pml_x, pml_z, geom, scalars,
            energy_ends, ckpt_every); c15/c35 may be None (isotropic / VTI).  With a damping-scheme
    geometry, pml_x is the [Z][pitch] damping taper and pml_z is None.
    """

    @staticmethod
    def forward(ctx, c11, c13, c33, c55, bx, bz, c15, c35, wav, pml_x, pml_z, geom, scalars, energy_ends,
                ckpt_every):
        ctx.set_materialize_grads(False)
        coef = geom.pack_coef([t.detach() if t is not None else None for t in (c11, c13, c33, c55, bx, bz, c15, c35)])
        src = wav.detach().to(torch.float32).contiguous()
        need_grad = any(ctx.needs_input_grad[:9])
        k = int(ckpt_every) if ckpt_every else checkpoint_interval(geom.nt)
        absorb = (None, None, pml_x) if geom.damping else (pml_x, pml_z, None)
        resident = resident_trace_segments(geom, k) if need_grad else 0
        res = run_forward(geom, coef, absorb[0], absorb[1], src, scalars, record=True, energy_ends=energy_ends,
                          ckpt_every=k if need_grad else 0, damp=absorb[2], resident=resident)
        ctx.traces = res["traces"] if need_grad else None
        rec = res["rec"] or [torch.zeros((geom.S, geom.nt, geom.R), device=geom.device) for _ in range(5)]
        ctx.geom, ctx.k, ctx.scalars = geom, k, scalars
        ctx.ckpt = res["ckpt"] if need_grad else None
        ctx.shapes = [None if t is None else tuple(t.shape) for t in (c11, c13, c33, c55, bx, bz, c15, c35)]
        ctx.save_for_backward(coef, src, pml_x, pml_z)
        en = [e.contiguous() for e in res["energy"]]
        ctx.mark_non_differentiable(*en)
        return tuple(rec) + tuple(en)

    @staticmethod
    def backward(ctx, *grads):
        traces, ctx.traces = ctx.traces, None
        if all(g is None for g in grads[:5]):
            return (None,) * 15
        from .elastic_adjoint import run_backward
        coef, src, pml_x, pml_z = ctx.saved_tensors
        absorb = (None, None, pml_x) if ctx.geom.damping else (pml_x, pml_z, None)
        gcoef, gsrc = run_backward(ctx.geom, coef, src, absorb[0], absorb[1], ctx.scalars, ctx.ckpt, ctx.k, grads[:5],
                                   ctx.needs_input_grad[:9], damp=absorb[2], traces=traces)
        ctx.ckpt = None
        out = []
        for i, shp in enumerate(ctx.shapes):
            if shp is None or not ctx.needs_input_grad[i] or gcoef[i] is None:
                out.append(None)
            else:
                out.append(gcoef[i][:shp[0], :shp[1]].contiguous())
        return tuple(out) + (gsrc if ctx.needs_input_grad[8] else None,) + (None,) * 6


def forward_kernel(nx: int, nz: int, dx: float, dz: float, nt: int, dt: float,
                   nabc: int, free_surface: bool,
                   src_x: torch.Tensor, src_z: torch.Tensor, src_n: int, src_v: torch.Tensor, MT: torch.Tensor,
                   rcv_x: torch.Tensor, rcv_z: torch.Tensor, rcv_n: int,
                   abc_type: str, bcx: torch.Tensor, bcz: torch.Tensor, damp,
                   lamu, lam, bx: torch.Tensor, bz: torch.Tensor, CC: List[torch.Tensor],
                   fd_order: int = 4, n_segments: int = 1,
                   device: torch.device = torch.device("cpu"), dtype=torch.float32,
                   geometry: Optional[ElasticGeometry] = None, ckpt_every: int = 0,
                   src_fz: Optional[torch.Tensor] = None) -> Dict[str, torch.Tensor]:
    """Elastic forward modelling with the reference's signature and result dictionary
    (elastic_kernels.py:917-1037): the split-field PML scheme for abc_type 'pml' (any case, as the
    reference's driver tests it), the multiplicative-damping scheme with ``damp`` otherwise; at
    fd_order 4, or the three-term schemes of fd_order 6, 8 and 10 (step_forward_PML_6order,
    :426-584, and step_forward_ABL_6order, :781-912).  ``n_segments`` keeps its meaning for the energy maps (summed over the last step of
    every time chunk); memory is governed by the sqrt(nt) wavefield checkpoints of the HIP
    adjoint.  ``src_fz`` (extension, no counterpart in the reference): per-shot vertical-force
    weights f; the shot's wavelet w then also enters vz at the source point right after the
    velocity update (f w / 2 into each PML split, f w in the damping scheme)."""
    nat.require_gpu(device)
    if dtype != torch.float32:
        raise NotImplementedError("the HIP propagator computes in float32 (dtype=torch.float32)")
    use_pml = abc_type.lower() == "pml"
    if fd_order not in (4, 6, 8, 10):
        raise NotImplementedError(f"the HIP elastic propagator implements fd_order 4, 6, 8 and 10 (got {fd_order})")
    if use_pml and (bcx is None or bcz is None):
        raise ValueError("the PML scheme needs the bcx / bcz damping profiles")
    if not use_pml and damp is None:
        raise ValueError(f"abc_type {abc_type!r} needs the damping taper 'damp'")
    fso = fd_order // 2
    X = nx + 2 * nabc
    Z = nz + (nabc + fso if free_surface else 2 * nabc + fso)
    pad = lambda a: pad_torchSingle(a.to(device), nabc, fso, free_surface, device)
    c11, c13, c15, c33, c35, c55 = (pad(CC[i]) for i in (0, 2, 4, 11, 13, 18))
    bxp, bzp = pad(bx), pad(bz)
    aniso = bool((CC[4] != 0).any() or (CC[13] != 0).any())
    fdc = DiffCoef(fso, "s")
    if geometry is None or geometry.damping == use_pml or geometry.nn != fso:
        zoff = fso if free_surface else nabc + fso
        MTd = MT.to(device=device, dtype=torch.float32)
        if use_pml:
            half = MTd / 2
            src_m = torch.stack([-half[:, 0, 0], -half[:, 2, 2], -half[:, 0, 2]], dim=1)
        else:
            src_m = torch.stack([(-1 / 3) * MTd[:, 0, 0], (-1 / 3) * MTd[:, 2, 2], (-1 / 3) * MTd[:, 0, 2]], dim=1)
        geometry = ElasticGeometry(Z, X, nt, free_surface, (src_z + zoff).cpu().numpy(),
                                   (src_x + nabc).cpu().numpy(), src_m, (rcv_z + zoff).cpu().numpy(),
                                   (rcv_x + nabc).cpu().numpy(), device, aniso=aniso, damping=not use_pml,
                                   src_fz=src_fz, nn=fso, c2=float(fdc[2]) if fso > 2 else 0.0)
    elif geometry.aniso != aniso:
        geometry.aniso = aniso
    if use_pml:
        bcx_p = pad_torchSingle(bcx.to(device=device, dtype=torch.float32), 0, fso, free_surface, device)
        bcz_p = pad_torchSingle(bcz.to(device=device, dtype=torch.float32), 0, fso, free_surface, device)
        absorb_a, absorb_b = pml_factors(bcx_p, bcz_p, dt, geometry.pitch)
    else:
        dmp = pad_torchSingle(torch.as_tensor(damp).to(device=device, dtype=torch.float32), 0, fso, free_surface,
                              device)
        absorb_a, absorb_b = geometry.pack_plane(dmp), None
    f32 = lambda v: float(np.float32(v))
    scalars = (f32(dt), f32(dt / dx), f32(dt / dz), f32(dx), f32(dz), float(fdc[0]), float(fdc[1]))
    if src_v.dim() == 1:
        src_v = src_v.unsqueeze(0).expand(src_n, -1)
    ends = chunk_ends(nt, n_segments)
    outs = ElasticWave.apply(c11, c13, c33, c55, bxp, bzp, c15 if aniso else None, c35 if aniso else None,
                             src_v, absorb_a, absorb_b, geometry, scalars, ends, ckpt_every)
    row0 = fso if free_surface else fso + nabc
    res = {n: outs[i] for i, n in enumerate(NAMES)}
    for i, n in enumerate(NAMES):
        res["forward_wavefield_" + n] = outs[5 + i][row0:row0 + nz, nabc:nabc + nx]
    return res
