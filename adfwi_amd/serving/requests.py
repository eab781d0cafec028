"""Modelling and gradient requests: the seeded request stream, each request's problem, and its
HBM footprint.

A request asks for acoustic or isotropic-elastic modelling ("forward": the records) or for a
gradient ("gradient": d/d model of 0.5 * sum(records^2) over the recorded components, an
adjoint source that needs no observed data).  Its model comes from its own seed through the
recipes of ``adfwi_amd.workloads`` (the acoustic benchmark recipe, the elastic layered recipe),
scaled to the request's grid -- or, when the request names a shared model (`model`), from that
model's seed, so that requests against one model (the shots of one FWI iteration) can share
launches: see `pack_key` and `execute_pack`.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ..propagator.acoustic_kernels import PITCH_ALIGN, checkpoint_interval

KINDS = ("acoustic", "elastic")
MODES = ("forward", "gradient")
# an acoustic gradient request keeps every step's stencil images from its forward pass (no
# checkpoints, no recompute) when they take at most this many bytes; larger requests keep
# sqrt(nt) checkpoints and recompute each segment's images in the backward
RESIDENT_IMAGE_CAP_BYTES = 32 << 30
# a gradient request whose footprint would exceed this runs its shots in equal chunks, one
# forward and backward per chunk, so that two of the largest requests fit one GPU together
REQUEST_CAP_BYTES = 128 << 30


def _padded_shots(req: "Request", shots: Optional[int] = None) -> int:
    s = req.nshots if shots is None else int(shots)
    return -(-s // 4) * 4 if req.kind == "acoustic" else s


def images_resident(req: "Request", shots: Optional[int] = None) -> bool:
    """Whether an acoustic gradient launch of `shots` shots (default: the request's, in its shot
    chunks) keeps all its stencil images resident."""
    Z, X = req.padded
    pitch = -(-X // PITCH_ALIGN) * PITCH_ALIGN
    if shots is None:
        shots = shot_chunk(req)
    return req.nt * _padded_shots(req, shots) * 4 * Z * pitch <= RESIDENT_IMAGE_CAP_BYTES


def shot_chunk(req: "Request") -> int:
    """Shots per launch: all of the request's, except for a gradient request whose footprint
    would exceed REQUEST_CAP_BYTES; that one runs in chunks of at most this many shots (halved
    until one chunk fits), its gradients summed over the chunks."""
    n = req.nshots
    if req.mode == "gradient":
        while n > 1 and _footprint(req, n) > REQUEST_CAP_BYTES:
            n = -(-n // 2)
    return n


@dataclass
class Request:
    rid: int
    kind: str                 # "acoustic" | "elastic"
    mode: str                 # "forward" | "gradient"
    nz: int
    nx: int
    nt: int
    nshots: int
    seed: int
    arrival: float = 0.0      # seconds after the start of the stream
    nabc: int = 20
    keep: bool = False        # keep the full result (otherwise a digest)
    model: Optional[int] = None   # shared model key (its seed); None: the request's own model
    extra: dict = field(default_factory=dict)

    @property
    def model_seed(self) -> int:
        return self.seed if self.model is None else int(self.model)

    @property
    def dx(self) -> float:
        return 12.5 if self.kind == "acoustic" else 5.0

    @property
    def dt(self) -> float:
        return 1e-3 if self.kind == "acoustic" else 5e-4

    @property
    def f0(self) -> float:
        return 8.0

    @property
    def free_surface(self) -> bool:
        return self.kind == "elastic"

    @property
    def padded(self):
        if self.kind == "acoustic":
            return self.nz + 2 * self.nabc, self.nx + 2 * self.nabc
        return self.nz + self.nabc + 2, self.nx + 2 * self.nabc


# service-time model of one request on one MI355X: padded points x nt x shots at the measured
# rate of each kind and mode (the 4th-order kernels of the benchmark workloads, rounded)
_PT_STEPS_PER_S = {("acoustic", "forward"): 230e9, ("acoustic", "gradient"): 66e9,
                   ("elastic", "forward"): 30e9, ("elastic", "gradient"): 10e9}


def service_estimate_s(req: Request) -> float:
    """Seconds one request takes alone on the device (a rough model, for scheduling)."""
    Z, X = req.padded
    return Z * X * req.nt * req.nshots / _PT_STEPS_PER_S[(req.kind, req.mode)]


def footprint_bytes(req: Request) -> int:
    """Upper bound of the device memory one request holds while it runs: state and work sets,
    sqrt(nt) checkpoints, the segment images (acoustic) or trace (elastic), records, their
    adjoints, coefficient planes, gradient accumulators and model-side tensors.  A gradient
    request too large for one launch counts one of its shot chunks (`shot_chunk`)."""
    return _footprint(req, shot_chunk(req))


def _footprint(req: Request, shots: int) -> int:
    """footprint_bytes of one launch of `shots` of the request's shots."""
    Z, X = req.padded
    pitch = -(-X // PITCH_ALIGN) * PITCH_ALIGN
    plane = 4 * Z * pitch
    S, nt, R = _padded_shots(req, shots), req.nt, req.nx   # acoustic: padded to a coefficient group
    k = checkpoint_interval(nt)
    nseg = -(-nt // k)
    grad = req.mode == "gradient"
    if req.kind == "acoustic":
        nf, ncomp, ncoef = 3, 3, 5
    else:
        nf, ncomp, ncoef = 10, 5, 8
    state = nf * S * plane
    b = 2 * state                                   # forward ping-pong
    b += ncomp * S * nt * R * 4                     # records
    b += (ncoef + 16) * plane                       # coefficient planes and model-side fields
    b += 3 * 4 * S * nt                             # wavelets and their gradient
    if grad:
        b += 2 * state                              # adjoint ping-pong
        if req.kind == "acoustic":
            if images_resident(req, shots):
                # every segment's stencil images stay resident from the forward pass (no
                # checkpoints, no recompute)
                b += nt * S * plane
            else:
                b += max(nseg - 1, 1) * state       # checkpoints
                b += 2 * state + 2 * k * S * plane  # recompute pair, stencil and pressure images
            b += 2 * S * plane                      # gradient accumulators (<= one plane per shot)
        else:
            b += max(nseg - 1, 1) * state           # checkpoints
            b += k * state + state                  # segment trace, zero state
            b += 4 * S * ncoef * plane              # gradient accumulators (<= one set per shot)
        b += ncomp * S * nt * R * 4                 # record adjoints
    return int(b * 1.05) + (64 << 20)               # allocator rounding, small tensors


def make_stream(n: int = 2000, rate: float = 1.0, seed: int = 4, *, acoustic_frac: float = 0.6,
                gradient_frac: float = 0.5, nz_range=(128, 512), nx_range=(256, 2048), nt_range=(1000, 6000),
                shots_range=(1, 64), nabc: int = 20, model_pool: int = 0, shape_pool: int = 0) -> List[Request]:
    """Seeded Poisson stream: kind 60/40 acoustic/elastic, mode 50/50 forward/gradient, nz and
    nx uniform over multiples of 32 in their ranges, nt and shot count uniform.

    model_pool > 0: every request is drawn against one of that many shared models (kind, grid
    and nt fixed per model; mode and shot count per request), as the shot requests of an
    inversion are; requests of one model and mode can then share launches.
    shape_pool > 0: every request takes one of that many (kind, grid, nt) shapes but its own
    model, as requests of many users on standard survey sizes do; acoustic modelling requests
    of one shape can then share launches."""
    rng = np.random.default_rng(seed)
    pool = []
    for m in range(model_pool or shape_pool):
        kind = "acoustic" if rng.random() < acoustic_frac else "elastic"
        nz = int(rng.choice(np.arange(nz_range[0], nz_range[1] + 1, 32)))
        nx = int(rng.choice(np.arange(nx_range[0], nx_range[1] + 1, 32)))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        pool.append((kind, nz, nx, nt, int(rng.integers(1 << 30))))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        if pool:
            kind, nz, nx, nt, mseed = pool[int(rng.integers(len(pool)))]
            if not model_pool:
                mseed = None
        else:
            kind = "acoustic" if rng.random() < acoustic_frac else "elastic"
            nz = int(rng.choice(np.arange(nz_range[0], nz_range[1] + 1, 32)))
            nx = int(rng.choice(np.arange(nx_range[0], nx_range[1] + 1, 32)))
            nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
            mseed = None
        mode = "gradient" if rng.random() < gradient_frac else "forward"
        ns = int(rng.integers(shots_range[0], shots_range[1] + 1))
        out.append(Request(i, kind, mode, nz, nx, nt, ns, int(rng.integers(1 << 30)), t, nabc, model=mseed))
    return out


def pack_key(req: Request, gradients_across_models: bool = False):
    """Requests with equal keys can share launches.  Acoustic modelling requests need only one
    grid and nt: each member's model becomes one coefficient set of the launch, indexed per shot
    (`_packed_acoustic_forward`).  Acoustic gradient requests pack against one model, or with
    `gradients_across_models` against any models of one grid and nt: the adjoint keeps one
    gradient plane per shot group, so that every member's gradient is summed from its own groups
    (`_packed_acoustic_gradient`).  Elastic modelling requests pack against one model (shots are
    independent); elastic gradients, and gradient requests that run in shot chunks, run alone.
    None: the request runs by itself."""
    if req.mode == "gradient" and shot_chunk(req) < req.nshots:
        return None
    if req.kind == "acoustic" and (req.mode == "forward" or gradients_across_models):
        return (req.kind, req.mode, req.nz, req.nx, req.nt, req.nabc)
    if req.model is None or (req.mode == "gradient" and req.kind != "acoustic"):
        return None
    return (req.kind, req.mode, req.nz, req.nx, req.nt, req.nabc, int(req.model))


# ---------------------------------------------------------------------------
# problems (GPU side)
# ---------------------------------------------------------------------------
def _acoustic_vp(nz, nx, rng):
    from ..workloads import _grf
    zz, xx = np.mgrid[0:nz, 0:nx]
    vp = 1500.0 + 3000.0 * zz / max(nz - 1, 1)
    vp = vp + _grf(rng, nz, nx, min(20, nz // 6 + 1), 300.0)
    dip = np.deg2rad(rng.uniform(5, 25))
    surf = rng.uniform(0.3, 0.7) * nz + np.tan(dip) * (xx - nx / 2)
    vp = vp + np.where(zz > surf, 250.0, 0.0)
    return np.clip(vp, 1500, 4700).astype(np.float32)


def shot_positions(req: Request):
    """(x, z) grid indices of the request's sources: evenly spaced, 2 cells deep (4 elastic)."""
    sx = np.linspace(min(4, req.nx - 1), max(req.nx - 5, 0), req.nshots).astype(int)
    sz = np.full(req.nshots, 2 if req.kind == "acoustic" else 4)
    return sx, sz


def build_problem(req: Request, device, shots=None, graph_steps: int = 0):
    """(propagator, model) of a request; its sources (`shot_positions`, or the given (x, z)
    arrays), receivers on every column 2 cells deep.  graph_steps > 0: acoustic modelling
    replays HIP graphs of that many steps (`adfwi_amd.propagator.graphs`)."""
    import torch
    from ..model import AcousticModel, IsotropicElasticModel
    from ..propagator import AcousticPropagator, ElasticPropagator
    from ..survey import Receiver, Source, Survey
    from ..workloads import elastic_layers, gardner, ricker
    rng = np.random.default_rng(req.model_seed)
    nz, nx = req.nz, req.nx
    src = Source(req.nt, req.dt, req.f0)
    sx, sz = shot_positions(req) if shots is None else shots
    src.add_sources(sx, sz, ricker(req.nt, req.dt, req.f0))
    rcv = Receiver(req.nt, req.dt)
    rcv.add_receivers(np.arange(nx), np.full(nx, 2), "pr" if req.kind == "acoustic" else "vz")
    survey = Survey(src, rcv)
    grad = req.mode == "gradient"
    if req.kind == "acoustic":
        vp = _acoustic_vp(nz, nx, rng)
        model = AcousticModel(0, 0, nx, nz, req.dx, req.dx, vp=vp, rho=gardner(vp), vp_grad=grad,
                              nabc=req.nabc, free_surface=False, device=device)
        prop = AcousticPropagator(model, survey, device=device)
        prop.graph_steps = int(graph_steps)
        # gradients keep every segment's stencil images when they fit under the cap (no
        # recompute), otherwise none (checkpoints and recompute): fixed amounts that
        # footprint_bytes counts
        prop.resident_images = (1 << 30) if images_resident(req) else 0
    else:
        vp, vs = elastic_layers(rng, nz, nx, 6, 1800, 4500, 200.0)
        model = IsotropicElasticModel(0, 0, nx, nz, req.dx, req.dx, vp=vp, vs=vs, rho=gardner(vp),
                                      vp_grad=grad, vs_grad=grad, rho_grad=grad, free_surface=True,
                                      nabc=req.nabc, auto_update_rho=False, device=device)
        prop = ElasticPropagator(model, survey, device=device)
        prop.resident_images = 0
    return prop, model


def execute(req: Request, device, graph_steps: int = 0):
    """Runs one request on the current stream; returns (tensors of the result, digest dict).
    graph_steps > 0: an acoustic modelling request replays HIP graphs of that many steps."""
    import torch
    prop, model = build_problem(req, device, graph_steps=graph_steps if req.mode == "forward" else 0)
    comps = ("p",) if req.kind == "acoustic" else ("vx", "vz")
    if req.mode == "forward":
        with torch.no_grad():
            rec = prop.forward()
        out = {k: rec[k] for k in comps}
    else:
        chunk = shot_chunk(req)
        loss = 0.0
        for s0 in range(0, req.nshots, chunk):     # one launch, or equal shot chunks
            shots = None if chunk == req.nshots else np.arange(s0, min(s0 + chunk, req.nshots))
            rec = prop.forward(shot_index=shots)
            part = 0.5 * sum(rec[k].square().sum() for k in comps)
            part.backward()                        # the parameters' .grad sum over the chunks
            loss = loss + part.detach()
            del rec, part
        out = {n: p.grad for n, p in model.named_parameters() if p.grad is not None}
        out["loss"] = loss.reshape(1)
    return out, _digest(out)


def _digest(out):
    return {k: (float(v.double().sum()), float(v.double().square().sum())) for k, v in out.items()}


def _packed_acoustic_gradient(reqs: List[Request], device):
    """Gradients of acoustic requests of one grid and nt in shared launches.

    Every member's shots are padded with silent shots (zero wavelet: their wavefield, records,
    adjoint and gradient terms are exactly zero) to a multiple of the adjoint's gradient group
    (of the coefficient group when the members have different models, each model then being one
    coefficient set indexed per shot), so that each accumulator plane holds one member's shots
    only; a member's coefficient gradient is the sum of its own planes, carried to vp through its
    model's coefficient construction.  Returns None when the layout does not give aligned groups
    (the members then run alone)."""
    import torch
    from ..propagator.acoustic_kernels import (AcousticGeometry, checkpoint_interval, coefficient_fields,
                                               run_backward, run_forward)
    r0 = reqs[0]
    shared = len({r.model_seed for r in reqs}) == 1
    sxs, szs, real = [], [], []
    group = 2
    for attempt in range(2):
        sxs, szs, real, offsets, sets = [], [], [], [], []
        for i, r in enumerate(reqs):
            sx, sz = shot_positions(r)
            pad = (-len(sx)) % group
            offsets.append(len(real))
            sxs += list(sx) + [int(sx[0])] * pad
            szs += list(sz) + [int(sz[0])] * pad
            real += [True] * len(sx) + [False] * pad
            sets += [i] * (len(sx) + pad)
        prop, model = build_problem(r0, device, shots=(np.array(sxs), np.array(szs)))
        nabc = r0.nabc
        geom = AcousticGeometry(prop.nz + 2 * nabc, prop.nx + 2 * nabc, r0.nt, False, nabc,
                                np.array(szs) + nabc, np.array(sxs) + nabc,
                                prop.rcv_z.cpu().numpy() + nabc, prop.rcv_x.cpu().numpy() + nabc, device)
        g = geom.gradient_group if shared else geom.coef_group
        if group % g == 0 and group % geom.gradient_group == 0:
            break
        group = g
    else:
        return None
    mask = torch.tensor(real, device=device, dtype=torch.float32).unsqueeze(1)
    if shared:
        model.forward()
        a1, k1, a2, k2, k3 = coefficient_fields(model.vp, model.rho, prop.damp, r0.nz, r0.nx, nabc, r0.dt, r0.dx,
                                                False, device)
        coef = geom.pack_coef(1.0 - k1, a1.detach(), 1.0 - k2, a2.detach(), 1.0 - k3)
        src = (r0.dt * prop.wavelet.to(torch.float32)) * mask
        members = [(a1, model)] * len(reqs)
    else:
        geom.set_coef_sets(sets)
        coefs, members = [], []
        src = torch.zeros((geom.S, r0.nt), device=device, dtype=torch.float32)
        for r, o in zip(reqs, offsets):
            p_r, m_r = build_problem(r, device)
            m_r.forward()
            a1, k1, a2, k2, k3 = coefficient_fields(m_r.vp, m_r.rho, p_r.damp, r.nz, r.nx, nabc, r.dt, r.dx,
                                                    False, device)
            coefs.append(geom.pack_coef(1.0 - k1, a1.detach(), 1.0 - k2, a2.detach(), 1.0 - k3))
            src[o:o + r.nshots] = r.dt * p_r.wavelet.to(torch.float32)
            members.append((a1, m_r))
        coef = torch.stack(coefs).contiguous()
    k = checkpoint_interval(r0.nt)
    # the members' footprints count their images when every member keeps them, else checkpoints
    resident = -(-r0.nt // k) if all(images_resident(r) for r in reqs) else 0
    res = run_forward(geom, coef, src.contiguous(), record=True, energy_u_from=None, ckpt_every=k,
                      resident=resident, need_a1=True, need_a2=False)
    rec_p = res["rec"][0]
    planes, _, _ = run_backward(geom, coef, src.contiguous(), res["ckpt"], k, (rec_p, None, None), True, False,
                                False, group_planes=True, images=res["images"])
    g = geom.gradient_group
    results = []
    for r, o, (a1, m_r) in zip(reqs, offsets, members):
        n = r.nshots + (-r.nshots) % group
        ga1 = planes[o // g:(o + n) // g].sum(0)
        grad, = torch.autograd.grad(a1, m_r.vp, ga1, retain_graph=True)
        rp = rec_p[o:o + r.nshots]
        out = {"vp": grad, "loss": (0.5 * rp.square().sum()).detach().reshape(1)}
        results.append((out, _digest(out)))
    return results


def _packed_acoustic_forward(reqs: List[Request], device):
    """Acoustic modelling requests of one grid and nt against different models in shared
    launches: each member's coefficient planes form one coefficient set, its shots (padded with
    silent shots to a multiple of the geometry's coefficient group, so that no kernel's shot
    group mixes models) index that set, and its records are split back out.  Records are
    bit-identical to the member's run alone (the same coefficient construction and kernels)."""
    import torch
    from ..propagator.acoustic_kernels import AcousticGeometry, coefficient_fields, run_forward
    r0 = reqs[0]
    nabc = r0.nabc
    probs = [build_problem(r, device) for r in reqs]
    prop0 = probs[0][0]
    group = 1
    for attempt in range(2):
        sxs, szs, sets, offsets = [], [], [], []
        for i, r in enumerate(reqs):
            sx, sz = shot_positions(r)
            pad = (-len(sx)) % group
            offsets.append(len(sxs))
            sxs += list(sx) + [int(sx[0])] * pad
            szs += list(sz) + [int(sz[0])] * pad
            sets += [i] * (len(sx) + pad)
        geom = AcousticGeometry(prop0.nz + 2 * nabc, prop0.nx + 2 * nabc, r0.nt, False, nabc,
                                np.array(szs) + nabc, np.array(sxs) + nabc, prop0.rcv_z.cpu().numpy() + nabc,
                                prop0.rcv_x.cpu().numpy() + nabc, device)
        if group % geom.coef_group == 0:
            break
        group = geom.coef_group
    else:
        return None
    geom.set_coef_sets(sets)
    coefs = []
    src = torch.zeros((geom.S, r0.nt), device=device, dtype=torch.float32)
    with torch.no_grad():
        for (prop, model), r, o in zip(probs, reqs, offsets):
            model.forward()
            a1, k1, a2, k2, k3 = coefficient_fields(model.vp, model.rho, prop.damp, r.nz, r.nx, nabc, r.dt, r.dx,
                                                    False, device)
            coefs.append(geom.pack_coef(1.0 - k1, a1, 1.0 - k2, a2, 1.0 - k3))
            src[o:o + r.nshots] = r.dt * prop.wavelet.to(torch.float32)
        res = run_forward(geom, torch.stack(coefs).contiguous(), src, record=True, energy_u_from=None)
    rec_p = res["rec"][0]
    results = []
    for r, o in zip(reqs, offsets):
        out = {"p": rec_p[o:o + r.nshots]}
        results.append((out, _digest(out)))
    return results


def execute_pack(reqs: List[Request], device, graph_steps: int = 0):
    """Runs requests with one `pack_key` in shared launches; returns [(result, digest)] in the
    members' order.  Each result is the member's own: its records (shots are independent), or
    its own gradient and loss.  graph_steps: as in `execute`, for modelling packs."""
    import torch
    if len(reqs) == 1:
        return [execute(reqs[0], device, graph_steps)]
    r0 = reqs[0]
    key = pack_key(r0, True)      # the widest packing the launches support
    if key is None or any(pack_key(r, True) != key for r in reqs):
        raise ValueError("requests of a pack must share one pack key")
    if r0.mode == "gradient":
        res = _packed_acoustic_gradient(reqs, device)
        return res if res is not None else [execute(r, device) for r in reqs]
    if r0.kind == "acoustic" and len({r.model_seed for r in reqs}) > 1:
        res = _packed_acoustic_forward(reqs, device)
        return res if res is not None else [execute(r, device, graph_steps) for r in reqs]
    pos = [shot_positions(r) for r in reqs]
    shots = (np.concatenate([p[0] for p in pos]), np.concatenate([p[1] for p in pos]))
    prop, _ = build_problem(r0, device, shots=shots, graph_steps=graph_steps)
    comps = ("p",) if r0.kind == "acoustic" else ("vx", "vz")
    with torch.no_grad():
        rec = prop.forward()
    results = []
    o = 0
    for r in reqs:
        out = {k: rec[k][o:o + r.nshots] for k in comps}
        results.append((out, _digest(out)))
        o += r.nshots
    return results
