"""Acoustic propagation on the MI355X: coefficient set-up in torch, time stepping in HIP.

Drop-in for ADFWI/propagator/acoustic_kernels.py: ``pad_torchSingle`` (:16-37)
and ``forward_kernel`` (:179-301) keep their signatures and results.  The
per-step op sequence (:40-176) and its autograd replay under
``torch.utils.checkpoint`` (:267-278) are replaced by ``AcousticWave``, a
``torch.autograd.Function`` whose forward runs the fused HIP step and whose
backward runs the hand-written discrete adjoint with sqrt(nt) wavefield
checkpoints and segment recomputation.
"""
import ctypes
import math
import os
import threading
import weakref
from typing import Dict, Optional

import numpy as np
import torch

from .. import _native as nat

C1 = 9.0 / 8.0
C2 = -1.0 / 24.0
PITCH_ALIGN = 64  # floats: rows start on 256-byte boundaries


def replicate_pad(a: torch.Tensor, left: int, right: int, top: int, bottom: int) -> torch.Tensor:
    """Edge-replicating pad of a 2-D array (the values of F.pad(mode='replicate')) whose backward
    sums each edge's pad cells with expand's reduction, which is deterministic, instead of the
    atomic adds of the replicate-pad backward, whose order varies from run to run."""
    nz, nx = a.shape
    parts = ([a[:1].expand(top, nx)] if top else []) + [a] + ([a[-1:].expand(bottom, nx)] if bottom else [])
    a = torch.cat(parts, 0) if len(parts) > 1 else a
    Z = a.shape[0]
    parts = ([a[:, :1].expand(Z, left)] if left else []) + [a] + ([a[:, -1:].expand(Z, right)] if right else [])
    return torch.cat(parts, 1) if len(parts) > 1 else a


def pad_torchSingle(v: torch.Tensor, pml: int, nz: int, nx: int, ns: int,
                    device: torch.device = torch.device("cpu")) -> torch.Tensor:
    """Edge-replicating pad (nz, nx) -> (nz+2*pml, nx+2*pml), float32 like the reference
    (acoustic_kernels.py:16-37 allocates the destination without a dtype)."""
    return replicate_pad(v.reshape(nz, nx).to(torch.float32), pml, pml, pml, pml)


def checkpoint_interval(nt: int) -> int:
    """Wavefield checkpoint spacing: ceil(sqrt(nt)) steps, so memory grows as sqrt(nt)."""
    return max(1, int(math.ceil(math.sqrt(nt))))


def last_chunk_start(nt: int, segments: int) -> int:
    """First step of the last torch.chunk(.., segments) time chunk (acoustic_kernels.py:268)."""
    segments = max(1, int(segments))
    size = -(-nt // segments)
    nchunks = -(-nt // size)
    return (nchunks - 1) * size


class Timeline:
    """Optional HIP-event timing of every library call (on the stream the kernels run on).

    When ``enabled``, each forward / recompute / adjoint call appends
    (label, kernel, launches, start_event, end_event); ``summary()`` turns them
    into total milliseconds and launches per kernel.
    """

    def __init__(self):
        self.enabled = False
        self.records = []

    def call(self, label, kernel, launches, device, fn, steps=None):
        """launches: kernel launches of the call; steps: time steps they advance (default: launches)."""
        if not self.enabled:
            return fn()
        st = torch.cuda.current_stream(device)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        out = fn()
        b.record(st)
        self.records.append((label, kernel, launches, launches if steps is None else steps, a, b))
        return out

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for label, kernel, n, nstep, a, b in self.records:
            d = out.setdefault(kernel, {"ms": 0.0, "launches": 0, "steps": 0, "calls": {}, "steps_by_call": {}})
            ms = a.elapsed_time(b)
            d["ms"] += ms
            d["launches"] += n
            d["steps"] += nstep
            d["calls"][label] = d["calls"].get(label, 0.0) + ms
            d["steps_by_call"][label] = d["steps_by_call"].get(label, 0) + nstep
        return out


def forward_launches(t0: int, n: int, nt: int, ckpt_every: int, fuse: bool):
    """Launches adfwi_acoustic_forward issues for steps t0 .. t0+n-1: (pairs, single steps).
    A step is paired with the next one unless it is the call's last step or its state is a
    checkpoint."""
    pairs = singles = 0
    k = 0
    while k < n:
        t = t0 + k
        ck = ckpt_every and (t + 1) % ckpt_every == 0 and t + 1 < nt
        if fuse and k + 1 < n and not ck:
            pairs += 1
            k += 2
        else:
            singles += 1
            k += 1
    return pairs, singles


TIMELINE = Timeline()


class AcousticGeometry:
    """Device-side description of one acoustic problem (grid, sources, receivers, tiling).

    Holds the int32 index tables and receiver tile bins; builds the ctypes
    descriptor for each call.  Padded indices are passed in.
    """

    def __init__(self, Z: int, X: int, nt: int, free_surface: bool, nabc: int,
                 src_z: np.ndarray, src_x: np.ndarray, rcv_z: np.ndarray, rcv_x: np.ndarray,
                 device, shots_per_block: int = 0, graph_steps: Optional[int] = None,
                 resident_images: Optional[int] = None):
        nat.require_gpu(device)
        self.device = torch.device(device)
        self.Z, self.X, self.nt = int(Z), int(X), int(nt)
        self.free_surface = bool(free_surface)
        self.fs = int(nabc) if free_surface else 1
        self.pitch = -(-self.X // PITCH_ALIGN) * PITCH_ALIGN
        src_z = np.asarray(src_z, dtype=np.int64).reshape(-1)
        src_x = np.asarray(src_x, dtype=np.int64).reshape(-1)
        rcv_z = np.asarray(rcv_z, dtype=np.int64).reshape(-1)
        rcv_x = np.asarray(rcv_x, dtype=np.int64).reshape(-1)
        self.S, self.R = len(src_z), len(rcv_z)
        zmin = self.fs if free_surface else 0
        if self.S < 1:
            raise ValueError("at least one shot is required")
        if np.any(src_z < zmin) or np.any(src_z > self.Z - 1) or np.any(src_x < 0) or np.any(src_x > self.X - 1):
            raise ValueError(f"source index outside the padded grid {self.Z} x {self.X}")
        if self.R and (np.any(rcv_z < 0) or np.any(rcv_z > self.Z - 2) or np.any(rcv_x < 0)
                       or np.any(rcv_x > self.X - 2)):
            raise ValueError(f"receiver index outside the padded grid {self.Z} x {self.X} "
                             "(u has X-1 columns and w has Z-1 rows)")
        if not shots_per_block:
            shots_per_block = int(os.environ.get("ADFWI_SHOTS_PER_BLOCK", "0"))
        self.shots_per_block = int(shots_per_block)
        self.adj_shots_per_block = int(os.environ.get("ADFWI_ADJ_SHOTS_PER_BLOCK", "0"))
        self.row_unroll = int(os.environ.get("ADFWI_ROW_UNROLL", "0"))
        self.tile_z = int(os.environ.get("ADFWI_TILE_Z", "0"))
        # forward runs without checkpoints replay HIP graphs of this many steps (0: direct launches)
        self.graph_steps = int(os.environ.get("ADFWI_GRAPH_STEPS", "0") if graph_steps is None else graph_steps)
        # checkpoint segments (the last ones) whose stencil images the forward pass keeps for the
        # backward instead of recomputing them there: None = as many as the memory budget holds
        # (resident_segments), 0 = none; ADFWI_RESIDENT_IMAGES sets it when the argument is None
        env = os.environ.get("ADFWI_RESIDENT_IMAGES")
        self.resident_images = (int(env) if env is not None and resident_images is None else resident_images)
        self.coef_set = None      # per-shot coefficient set (set_coef_sets); None: one set
        i32 = lambda a: torch.tensor(a.astype(np.int32), device=self.device)
        self.src_z, self.src_x = i32(src_z), i32(src_x)
        self.rcv_z, self.rcv_x = i32(rcv_z), i32(rcv_x)
        # two steps per launch (the intermediate step's receivers written from inside the kernel)
        self.fuse_steps = int(os.environ.get("ADFWI_FUSE_STEPS", "1"))
        self.rcv_map = self._receiver_map(rcv_z, rcv_x) if self.fuse_steps else None
        # what a captured launch sequence depends on besides its buffers and the source positions
        # (the kernels read those from device memory, so a replay can take new ones)
        self.content_key = (self.Z, self.X, self.nt, self.S, self.free_surface, self.fs, self.shots_per_block,
                            self.adj_shots_per_block, self.row_unroll, self.tile_z, self.fuse_steps,
                            rcv_z.tobytes(), rcv_x.tobytes())
        desc = self.desc(None, None)
        self.layout = nat.Layout()
        nat.check(nat.load().adfwi_acoustic_get_layout(ctypes.byref(desc), ctypes.byref(self.layout)),
                  "adfwi_acoustic_get_layout")

    @property
    def gradient_group(self) -> int:
        """Consecutive shots that share one gradient accumulator plane in the adjoint."""
        return max(1, min(int(self.layout.shots_per_block), int(self.layout.reserved)))

    def _receiver_map(self, rcv_z: np.ndarray, rcv_x: np.ndarray) -> Optional[torch.Tensor]:
        """[Z] receivers per row, then [Z][pitch] receiver index (-1 elsewhere); None when
        receivers repeat a point or sit on the free-surface mirror row fs-1 (records are then
        taken one step per launch)."""
        if self.R == 0:
            return None
        pts = rcv_z.astype(np.int64) * self.pitch + rcv_x.astype(np.int64)
        if len(np.unique(pts)) != len(pts) or (self.free_surface and np.any(rcv_z == self.fs - 1)):
            return None
        m = np.full(self.Z + self.Z * self.pitch, -1, dtype=np.int32)
        m[:self.Z] = np.bincount(rcv_z, minlength=self.Z)[:self.Z]
        m[self.Z + pts] = np.arange(self.R, dtype=np.int32)
        return torch.tensor(m, device=self.device)

    def desc(self, coef: Optional[torch.Tensor], src: Optional[torch.Tensor]) -> nat.AcousticDesc:
        d = nat.AcousticDesc()
        d.abi_version = nat.ABI_VERSION
        d.struct_size = ctypes.sizeof(nat.AcousticDesc)
        d.nshots, d.nz, d.nx, d.pitch = self.S, self.Z, self.X, self.pitch
        d.fs, d.free_surface, d.nt, d.nrec = self.fs, int(self.free_surface), self.nt, self.R
        d.shots_per_block = self.shots_per_block
        d.adj_shots_per_block = self.adj_shots_per_block
        d.row_unroll = self.row_unroll
        d.tile_z = self.tile_z
        d.coef = None if coef is None else coef.data_ptr()
        d.src = None if src is None else src.data_ptr()
        d.src_z, d.src_x = self.src_z.data_ptr(), self.src_x.data_ptr()
        d.rcv_z, d.rcv_x = self.rcv_z.data_ptr(), self.rcv_x.data_ptr()
        d.coef_set = None if self.coef_set is None else self.coef_set.data_ptr()
        return d

    def set_coef_sets(self, sets) -> None:
        """One coefficient-set index per shot (None: one set for all shots): the coefficient
        tensor then holds [nsets][5][Z][pitch].  Every aligned group of `coef_group` shots must
        use one set."""
        if sets is None:
            self.coef_set = None
            return
        sets = np.asarray(sets, dtype=np.int64).reshape(-1)
        if len(sets) != self.S:
            raise ValueError(f"coefficient sets: one per shot ({self.S}), got {len(sets)}")
        g = self.coef_group
        for s0 in range(0, self.S, g):
            if len(set(sets[s0:s0 + g].tolist())) != 1:
                raise ValueError(f"shots {s0}..{min(s0 + g, self.S) - 1} of one group of {g} use different "
                                 "coefficient sets; pad each member to a multiple of coef_group shots")
        self.coef_set = torch.tensor(sets.astype(np.int32), device=self.device)

    @property
    def coef_group(self) -> int:
        """Aligned shot groups within which the kernels load one coefficient set."""
        return max(1, int(self.layout.shots_per_block), int(self.layout.reserved))

    # -- buffers ----------------------------------------------------------
    def state(self, n: int = 1) -> torch.Tensor:
        return torch.zeros((n, 3, self.S, self.Z, self.pitch), device=self.device, dtype=torch.float32)

    def pack_coef(self, omk1, a1, omk2, a2, omk3) -> torch.Tensor:
        coef = torch.zeros((5, self.Z, self.pitch), device=self.device, dtype=torch.float32)
        for k, f in enumerate((omk1, a1, omk2, a2, omk3)):
            coef[k, :, :self.X] = f
        return coef


class ResidentSet(dict):
    """{segment: buffers} kept resident by a forward pass for its backward.  While it lives, the
    memory its backward will still allocate (recompute buffers, adjoint states) is set aside from
    the budget of later forward passes, so that several pending backward passes (losses summed
    over batches before one backward) cannot overcommit the device."""
    _lock = threading.Lock()
    pending_bytes = 0

    def hold(self, nbytes: int) -> "ResidentSet":
        nbytes = int(nbytes)
        with ResidentSet._lock:
            ResidentSet.pending_bytes += nbytes

        def release():
            with ResidentSet._lock:
                ResidentSet.pending_bytes -= nbytes
        weakref.finalize(self, release)
        return self


def resident_budget_bytes(device, other_bytes: int) -> int:
    """Device memory that segments kept resident for the backward may take: 80 % of the free
    memory (what torch's caching allocator holds unused included: the previous iteration's
    buffers, typically) less 8 GiB, less `other_bytes` (the checkpoints and the backward's own
    buffers still to be allocated) and less the reserves of pending backward passes
    (ResidentSet); ADFWI_IMAGE_BUDGET_BYTES overrides it."""
    budget = int(os.environ.get("ADFWI_IMAGE_BUDGET_BYTES", "0"))
    if not budget:
        free, _ = torch.cuda.mem_get_info(device)
        free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
        budget = max(0, free * 4 // 5 - (8 << 30) - int(other_bytes) - ResidentSet.pending_bytes)
    return budget


def resident_segments(geom: AcousticGeometry, k: int, need_a1: bool, need_a2: bool) -> int:
    """How many of the last checkpoint segments keep their stencil (and pressure) images from the
    forward pass, so that the backward skips their recomputation: at most geom.resident_images
    (0 = none); with None, as many as resident_budget_bytes allows once the checkpoints, one
    segment's recompute images and the four states of the backward are set aside."""
    planes = int(bool(need_a1)) + int(bool(need_a2))
    nseg = -(-geom.nt // k)
    cap = geom.resident_images
    if planes == 0 or cap == 0:
        return 0
    seg_bytes = planes * k * geom.S * geom.Z * geom.pitch * 4
    if cap is None:
        state_bytes = 3 * geom.S * geom.Z * geom.pitch * 4
        cap = resident_budget_bytes(geom.device, (nseg + 4) * state_bytes + seg_bytes) // seg_bytes
    return int(max(0, min(cap, nseg)))


def run_forward(geom: AcousticGeometry, coef: torch.Tensor, src: torch.Tensor, *, record: bool = True,
                energy_u_from: Optional[int] = None, ckpt_every: int = 0, resident: int = 0,
                need_a1: bool = True, need_a2: bool = False):
    """Runs all nt forward steps from a zero state.

    Returns dict(rec=[3 x (S, nt, R)] or None, energy=(3, Z, X) or None: sum over steps and shots of
    p^2, sum over steps >= energy_u_from of u^2, and sum over shots of w^2 after the last step;
    ckpt=(n, 3, S, Z, pitch) or None, images={segment: (image, image_p)}).

    resident > 0 (with ckpt_every): the last `resident` checkpoint segments are run with the
    stencil image (need_a1) and the pressure image (need_a2) written to buffers of their own,
    which run_backward then uses instead of recomputing those segments; their checkpoints are
    not written.
    """
    lib = nat.load()
    dev = geom.device
    S, Z, X, nt, R = geom.S, geom.Z, geom.X, geom.nt, geom.R
    L = geom.layout
    work = geom.state(2)
    ckpt = None
    nseg = -(-nt // ckpt_every) if ckpt_every else 1
    resident = int(resident) if ckpt_every and (need_a1 or need_a2) else 0
    head = nt if not resident else (nseg - resident) * ckpt_every    # steps before the resident segments
    if ckpt_every and head > 0:
        ckpt = geom.state(max(nseg - 1, 1) if not resident else max(nseg - resident, 1))
    rec = None
    if record and R:
        rec = [torch.zeros((S, nt, R), device=dev, dtype=torch.float32) for _ in range(3)]
    energy = None
    if energy_u_from is not None:
        energy = torch.zeros((L.ngroups, 2, Z, geom.pitch), device=dev, dtype=torch.float32)
    d = geom.desc(coef, src)
    fused = bool(geom.fuse_steps) and (rec is None or geom.rcv_map is not None)
    images = ResidentSet()
    if resident:
        planes = int(bool(need_a1)) + int(bool(need_a2))
        state_bytes = 3 * S * Z * geom.pitch * 4
        images.hold(4 * state_bytes + (planes * ckpt_every * S * Z * geom.pitch * 4 if resident < nseg else 0))

    def call(state_in, t0, nsteps, ckpt_buf, img, img_p, kind):
        a = nat.AcousticFwdArgs()
        a.state_in = state_in
        a.work[0], a.work[1] = work[0].data_ptr(), work[1].data_ptr()
        a.ckpt = None if ckpt_buf is None else ckpt_buf.data_ptr()
        a.ckpt_every = int(ckpt_every) if ckpt_buf is not None else 0
        a.t0, a.nsteps = t0, nsteps
        a.energy_u_from = int(energy_u_from) if energy_u_from is not None else nt
        if rec is not None:
            a.rec_p, a.rec_u, a.rec_w = rec[0].data_ptr(), rec[1].data_ptr(), rec[2].data_ptr()
        a.energy = None if energy is None else energy.data_ptr()
        a.image = None if img is None else img.data_ptr()
        a.image_p = None if img_p is None else img_p.data_ptr()
        a.rcv_map = None if geom.rcv_map is None else geom.rcv_map.data_ptr()
        a.fuse_steps = geom.fuse_steps
        pairs, singles = forward_launches(t0, nsteps, nt, int(a.ckpt_every), fused)
        TIMELINE.call(kind, "ac_fwd_pair" if fused else "ac_fwd_step", pairs + singles, dev, lambda: nat.check(
            lib.adfwi_acoustic_forward(ctypes.byref(d), ctypes.byref(a), nat.stream_handle(dev)),
            "adfwi_acoustic_forward"), steps=nsteps)
        return a.state_out

    state = work[0].data_ptr()          # zero initial state
    if head > 0:
        state = call(state, 0, head, ckpt, None, None, "forward")
    # The forward writes the stencil image at every point of the strips, rows [zbase, Z) x
    # columns [0, X), at every step; only the rows above the strips and the pitch padding need
    # clearing.  ADFWI_POISON_RESIDENT=1 fills the rest with NaN first (tests: a point the
    # adjoint read without the forward writing it would show in the gradient).
    zbase = int(geom.layout.zbase)
    poison = os.environ.get("ADFWI_POISON_RESIDENT") == "1"

    def image_buffer(shape):
        img = torch.empty(shape, device=dev, dtype=torch.float32)
        if poison:
            img.fill_(float("nan"))
        if zbase:
            img[:, :, :zbase].zero_()
        img[..., X:].zero_()
        return img

    for seg in range(nseg - resident, nseg):
        t_lo = seg * ckpt_every
        n = min(nt, t_lo + ckpt_every) - t_lo
        shape = (n, S, Z, geom.pitch)              # the last segment may be shorter
        img = image_buffer(shape) if need_a1 else None
        img_p = torch.zeros(shape, device=dev, dtype=torch.float32) if need_a2 else None
        state = call(state, t_lo, n, None, img, img_p, "forward+images")
        images[seg] = (img, img_p)
    out = {"rec": rec, "ckpt": ckpt, "energy": None, "images": images}
    if energy is not None:
        final = work[0] if state == work[0].data_ptr() else (work[1] if state == work[1].data_ptr() else None)
        if final is None:   # last step wrote a checkpoint slot (only when nt is a multiple of k, never)
            raise RuntimeError("final state not in a work buffer")
        e = energy.sum(0)[:, :, :X]
        e_w = final[2, :, :, :X].square().sum(0)
        out["energy"] = torch.stack([e[0], e[1], e_w])
    return out


def run_backward(geom: AcousticGeometry, coef: torch.Tensor, src: torch.Tensor, ckpt: Optional[torch.Tensor],
                 ckpt_every: int, g_rec, need_a1: bool, need_a2: bool, need_src: bool,
                 group_planes: bool = False, images: Optional[dict] = None, noalloc: bool = False):
    """Hand-written adjoint over all steps with segment recomputation from the checkpoints.

    g_rec: tuple of three (S, nt, R) adjoint sources (or None each).
    images: {segment: (image, image_p)} kept by run_forward(resident=...); those segments are not
    recomputed, and each buffer is released once its segment is done.
    Returns (grad_a1 (Z, X) or None, grad_a2 or None, grad_src (S, nt) or None).
    group_planes: grad_a1 / grad_a2 are returned unsummed, (G, Z, X) with one plane per group of
    `geom.gradient_group` consecutive shots (their sum is the gradient of all shots).
    noalloc: the adjoint runs through adfwi_acoustic_adjoint_noalloc, whose per-group planes live
    in the library's scratch array and are summed into one plane per gradient by each call.
    """
    if noalloc and group_planes:
        raise ValueError("noalloc sums the group planes inside each call; group_planes needs them apart")
    lib = nat.load()
    dev = geom.device
    S, Z, X, nt = geom.S, geom.Z, geom.X, geom.nt
    L = geom.layout
    k = ckpt_every
    nseg = -(-nt // k)
    rw = geom.state(2)    # recompute ping-pong
    lw = geom.state(2)    # adjoint ping-pong
    images = {} if images is None else images
    recompute_any = len(images) < nseg
    image = torch.zeros((k, S, Z, geom.pitch), device=dev, dtype=torch.float32) if need_a1 and recompute_any else None
    image_p = torch.zeros((k, S, Z, geom.pitch), device=dev, dtype=torch.float32) if need_a2 and recompute_any else None
    nacc = 1 if noalloc else int(L.accum_floats) // (Z * geom.pitch)   # accumulator planes (finest shot grouping)
    ga1 = torch.zeros((nacc, Z, geom.pitch), device=dev, dtype=torch.float32) if need_a1 else None
    ga2 = torch.zeros((nacc, Z, geom.pitch), device=dev, dtype=torch.float32) if need_a2 else None
    gsrc = torch.zeros((S, nt), device=dev, dtype=torch.float32) if need_src else None
    g = [None if x is None else x.contiguous().to(torch.float32) for x in g_rec]
    gp = [None if x is None else x.data_ptr() for x in g]
    d = geom.desc(coef, src)
    stream = nat.stream_handle(dev)
    lam = lw[0]   # zero adjoint state after the last step; its receiver sources are added in the call
    need_image = need_a1 or need_a2
    for seg in reversed(range(nseg)):
        t_lo = seg * k
        t_hi = min(nt, (seg + 1) * k) - 1
        n = t_hi - t_lo + 1
        kept = images.pop(seg, None)
        if kept is not None:
            seg_image, seg_image_p = kept
        else:
            seg_image, seg_image_p = image, image_p
        if need_image and kept is None:
            if seg == 0:
                rw[0].zero_()
                sin = rw[0]
            else:
                sin = ckpt[seg - 1]
            fa = nat.AcousticFwdArgs()
            fa.state_in = sin.data_ptr()
            fa.work[0], fa.work[1] = rw[0].data_ptr(), rw[1].data_ptr()
            fa.t0, fa.nsteps = t_lo, n
            fa.energy_u_from = nt
            fa.image = None if image is None else image.data_ptr()
            fa.image_p = None if image_p is None else image_p.data_ptr()
            fa.fuse_steps = geom.fuse_steps
            pairs, singles = forward_launches(t_lo, n, nt, 0, bool(geom.fuse_steps))
            TIMELINE.call("recompute", "ac_fwd_pair" if geom.fuse_steps else "ac_fwd_step", pairs + singles,
                          dev, lambda: nat.check(
                lib.adfwi_acoustic_forward(ctypes.byref(d), ctypes.byref(fa), stream),
                "adfwi_acoustic_forward (recompute)"), steps=n)
        aa = nat.AcousticAdjArgs()
        aa.lam_in = lam.data_ptr()
        aa.work[0], aa.work[1] = lw[0].data_ptr(), lw[1].data_ptr()
        aa.t_hi, aa.nsteps, aa.image_t0 = t_hi, n, t_lo
        aa.image = None if seg_image is None or not need_a1 else seg_image.data_ptr()
        aa.image_p = None if seg_image_p is None or not need_a2 else seg_image_p.data_ptr()
        aa.grad_rec_p, aa.grad_rec_u, aa.grad_rec_w = gp
        if noalloc:
            aa.grad_a1 = aa.grad_a2 = None
        else:
            aa.grad_a1 = None if ga1 is None else ga1.data_ptr()
            aa.grad_a2 = None if ga2 is None else ga2.data_ptr()
        aa.grad_src = None if gsrc is None else gsrc.data_ptr()
        aa.rcv_map = None if geom.rcv_map is None else geom.rcv_map.data_ptr()
        aa.fuse_steps = geom.fuse_steps
        fused = bool(geom.fuse_steps) and (all(x is None for x in g) or geom.rcv_map is not None)
        if noalloc:
            call = lambda: nat.check(lib.adfwi_acoustic_adjoint_noalloc(
                ctypes.byref(d), ctypes.byref(aa), nat.ptr(ga1), nat.ptr(ga2), stream),
                "adfwi_acoustic_adjoint_noalloc")
        else:
            call = lambda: nat.check(lib.adfwi_acoustic_adjoint(ctypes.byref(d), ctypes.byref(aa), stream),
                                     "adfwi_acoustic_adjoint")
        TIMELINE.call("adjoint", "ac_adj_pair" if fused else "ac_adj_step", (n + 1) // 2 if fused else n, dev,
                      call, steps=n)
        lam = lw[0] if aa.lam_out == lw[0].data_ptr() else lw[1]
        kept = seg_image = seg_image_p = None     # a kept segment's images are free for reuse now
    if group_planes:
        return (None if ga1 is None else ga1[:, :, :X], None if ga2 is None else ga2[:, :, :X], gsrc)
    out_a1 = ga1.sum(0)[:, :X] if ga1 is not None else None
    out_a2 = ga2.sum(0)[:, :X] if ga2 is not None else None
    return out_a1, out_a2, gsrc


class AcousticWave(torch.autograd.Function):
    """Records (p, u, w) and energy maps from the padded coefficient fields.

    forward(a1, a2, src, omk1, omk2, omk3, geom, energy_u_from, ckpt_every)
      a1 = rho c^2 dt/dz, a2 = dt/(rho dz), src = dt * wavelet (S, nt),
      omk* = 1 - kappa* damping factors (not differentiated).
    Returns rec_p, rec_u, rec_w (S, nt, R), e_p, e_u, e_w (Z, X; non-differentiable):
      e_p = sum_t sum_s p^2, e_u = sum_{t >= energy_u_from} sum_s u^2, e_w = sum_s w^2 at t = nt-1.
    """

    @staticmethod
    def forward(ctx, a1, a2, src, omk1, omk2, omk3, geom, energy_u_from, ckpt_every):
        ctx.set_materialize_grads(False)
        coef = geom.pack_coef(omk1, a1.detach(), omk2, a2.detach(), omk3)
        src = src.detach().to(torch.float32).contiguous()
        need_grad = any(ctx.needs_input_grad[:3])  # grad mode is off inside forward()
        k = int(ckpt_every) if ckpt_every else checkpoint_interval(geom.nt)
        need_a1, need_a2 = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not need_grad and geom.graph_steps and geom.R:
            from .graphs import graphed_forward
            res = graphed_forward(geom, coef, src, energy_u_from, geom.graph_steps)
        else:
            resident = resident_segments(geom, k, need_a1, need_a2) if need_grad else 0
            res = run_forward(geom, coef, src, record=True, energy_u_from=energy_u_from,
                              ckpt_every=k if need_grad else 0, resident=resident, need_a1=need_a1,
                              need_a2=need_a2)
        ctx.images = res.get("images") if need_grad else None
        rec = res["rec"]
        if rec is None:
            rec = [torch.zeros((geom.S, geom.nt, geom.R), device=geom.device) for _ in range(3)]
        e = res["energy"]
        ctx.geom, ctx.k = geom, k
        ctx.ckpt = res["ckpt"] if need_grad else None
        ctx.save_for_backward(coef, src)
        outs = (rec[0], rec[1], rec[2], e[0].contiguous(), e[1].contiguous(), e[2].contiguous())
        ctx.mark_non_differentiable(*outs[3:])
        return outs

    @staticmethod
    def backward(ctx, g_p, g_u, g_w, *_):
        coef, src = ctx.saved_tensors
        need_a1, need_a2, need_src = ctx.needs_input_grad[:3]
        images, ctx.images = ctx.images, None
        if g_p is None and g_u is None and g_w is None:
            return (None,) * 9
        ga1, ga2, gsrc = run_backward(ctx.geom, coef, src, ctx.ckpt, ctx.k, (g_p, g_u, g_w),
                                      need_a1, need_a2, need_src, images=images)
        ctx.ckpt = None
        return ga1, ga2, gsrc, None, None, None, None, None, None


def coefficient_fields(v, rho, damp, nz, nx, nabc, dt, dz, free_surface, device):
    """A1, K1, A2, K2, K3 exactly as acoustic_kernels.py:225-265 forms them (dz on both axes)."""
    c = pad_torchSingle(v, nabc, nz, nx, 0, device)
    den = pad_torchSingle(rho, nabc, nz, nx, 0, device)
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    fs = nabc if free_surface else 1
    # true division, as the reference's CPU arithmetic does: torch on a GPU turns division by a
    # Python number into multiplication by its reciprocal, which differs in the last bit
    alpha1 = den * c * c * dt / torch.full([], dz, dtype=den.dtype, device=den.device)
    kappa1 = damp * dt
    alpha2 = dt / (den * dz)
    kappa2 = torch.zeros_like(damp)
    kappa2[:, 1:X - 2] = 0.5 * (damp[:, 1:X - 2] + damp[:, 2:X - 1]) * dt
    kappa3 = torch.zeros_like(damp)
    kappa3[fs:Z - 2, :] = 0.5 * (damp[fs:Z - 2, :] + damp[fs + 1:Z - 1, :]) * dt
    return alpha1, kappa1, alpha2, kappa2, kappa3


def forward_kernel(nx: int, nz: int, dx: float, dz: float, nt: int, dt: float,
                   nabc: int, free_surface: bool,
                   src_x: torch.Tensor, src_z: torch.Tensor, src_n: int, src_v: torch.Tensor,
                   rcv_x: torch.Tensor, rcv_z: torch.Tensor, rcv_n: int,
                   damp: torch.Tensor, v: torch.Tensor, rho: torch.Tensor,
                   checkpoint_segments: int = 1,
                   device: torch.device = torch.device("cpu"), dtype: torch.dtype = torch.float32,
                   geometry: Optional[AcousticGeometry] = None, ckpt_every: int = 0) -> Dict[str, torch.Tensor]:
    """Acoustic forward modelling with the reference's signature and result dictionary
    (acoustic_kernels.py:179-301).

    ``checkpoint_segments`` keeps the reference meaning for the energy maps
    (the u/w maps depend on the last time chunk, acoustic_kernels.py:286-288);
    memory is governed by the sqrt(nt) wavefield checkpoints of the HIP adjoint
    (``ckpt_every`` overrides the interval).
    """
    nat.require_gpu(device)
    if dtype != torch.float32:
        raise NotImplementedError("the HIP propagator computes in float32 (dtype=torch.float32)")
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    a1, k1, a2, k2, k3 = coefficient_fields(v, rho, damp.to(device=device, dtype=torch.float32),
                                            nz, nx, nabc, dt, dz, free_surface, device)
    if src_v.dim() == 1:
        src_v = src_v.unsqueeze(0).expand(src_n, -1)
    src = dt * src_v
    if geometry is None:
        geometry = AcousticGeometry(Z, X, nt, free_surface, nabc,
                                    (src_z + nabc).cpu().numpy(), (src_x + nabc).cpu().numpy(),
                                    (rcv_z + nabc).cpu().numpy(), (rcv_x + nabc).cpu().numpy(), device)
    eu_from = last_chunk_start(nt, checkpoint_segments)
    rp, ru, rw, ep, eu, ew = AcousticWave.apply(a1, a2, src, 1.0 - k1, 1.0 - k2, 1.0 - k3, geometry,
                                                eu_from, ckpt_every)
    crop = (slice(nabc, nabc + nz), slice(nabc, nabc + nx))
    f_p = ep[crop]
    f_u = f_p + eu[crop]
    f_w = f_p + (eu[crop] + ew[crop])
    return {"p": rp, "u": ru, "w": rw,
            "forward_wavefield_p": f_p, "forward_wavefield_u": f_u, "forward_wavefield_w": f_w}
