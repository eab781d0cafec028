"""Parity at the benchmark workloads' own sizes, with each result's error against an FP64 run.

* 100 x 200 acoustic, 10 shots, nt 1000: the HIP path against the CPU oracle (fp32 and fp64).
* 384 x 1152 acoustic, 64 shots, nt 4000, sqrt(nt) checkpoints with resident segments (the
  benchmark's configuration): against the oracle on the GPU with 64 checkpoint segments (fp32),
  an FP64 oracle run on every eighth shot, and a central finite difference along a seeded
  smooth direction.
* 200 x 600 isotropic elastic (free surface, nt 4000) and 320 x 960 VTI elastic (nt 5000): the
  full grids and nt with explosive moment-tensor shots, against the oracle on the GPU (fp32 and
  fp64, 64 checkpoint segments), plus a central finite difference in the six coefficient fields.
* 20 requests of the seeded serving stream (grids 128-512 x 256-2048, nt 1000-6000, 1-64 shots,
  60/40 acoustic/elastic, 50/50 modelling/gradient): each result in the mix equals its run alone.

Bars: records rel-L2 <= 1e-5 and gradients <= 1e-4 against the fp32 oracle (the reference's
algorithm and op order); finite differences to 5e-3; the FP64 errors are recorded, and bounded
loosely (<= 1e-3) only to catch a gross fault.  The loss is a misfit against data modelled from
the true model: on the 100 x 200 workload the reference's waveform L2 (ADFWI/fwi/misfit/L2.py:
26-29, a square root per trace); on the larger ones the least-squares objective 1/2 dt
sum (obs - syn)^2, because there far-offset traces receive no arrival within the record, their
residual is exactly zero and the per-trace square root's gradient is NaN (in the reference too).
Every test prints one `PARITY {json}` row per compared result and appends it to the file named
by ADFWI_PARITY_LOG when that is set.
"""
import json
import os
import time

import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
REC_BAR, GRAD_BAR, FD_BAR, FP64_LOOSE = 1e-5, 1e-4, 5e-3, 1e-3


def _rel(a, b):
    """rel-L2 of a against b, in float64 on the tensors' device."""
    a = torch.as_tensor(a).double()
    b = torch.as_tensor(b).to(a.device).double()
    nb = float(b.norm())
    return float((a - b).norm()) / (nb if nb > 0 else 1.0)


def _log(config, result, **errs):
    row = dict(config=config, result=result,
               **{k: (v if v is None or isinstance(v, str) else float(v)) for k, v in errs.items()})
    print("PARITY " + json.dumps(row), flush=True)
    path = os.environ.get("ADFWI_PARITY_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(row) + "\n")


def _direction(seed, shape, sigma=20.0):
    d = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma, mode="reflect")
    return d / np.abs(d).max()


def _lsq_f64(obs, syn, dt):
    """The least-squares misfit 1/2 sum (obs - syn)^2 dt, evaluated in float64."""
    r = obs.double() - syn.double()
    return float(0.5 * (r * r * dt).sum())


def _misfit(product, kind, dt):
    """The misfit object: kind 'l2' (the reference's per-trace root) or 'lsq' (least squares),
    the product's kernels or the oracle's expressions."""
    if product:
        from adfwi_amd.fwi import Misfit_waveform_L2, Misfit_waveform_L2_squared
        return (Misfit_waveform_L2 if kind == "l2" else Misfit_waveform_L2_squared)(dt).forward
    from oracle import misfit as om
    f = om.waveform_l2 if kind == "l2" else om.waveform_l2_squared
    return lambda obs, syn: f(obs, syn, dt)


# ----------------------------------------------------------------------------- acoustic

def _acoustic(mod, dev, wl, vp, rho, damp, obs, shots, dtype=torch.float32, segments=1, grad=True, loss="lsq"):
    """forward_kernel of `mod` (product or oracle) for `shots`; misfit `loss` against obs[shots];
    returns (records p/u/w, forward_wavefield_p, vp gradient or None)."""
    kw = {} if mod.__name__.startswith("adfwi_amd") else {"dtype": dtype}
    v = torch.tensor(vp, dtype=dtype, device=dev, requires_grad=grad)
    r = torch.tensor(rho, dtype=dtype, device=dev)
    sel = np.asarray(shots)
    res = mod.forward_kernel(wl.nx, wl.nz, wl.dx, wl.dx, wl.nt, wl.dt, wl.nabc, wl.free_surface,
                             torch.tensor(wl.src_x[sel], device=dev), torch.tensor(wl.src_z[sel], device=dev),
                             len(sel), torch.tensor(wl.wavelets()[sel], dtype=dtype, device=dev),
                             torch.tensor(wl.rcv_x, device=dev), torch.tensor(wl.rcv_z, device=dev), len(wl.rcv_x),
                             damp.to(device=dev, dtype=dtype), v, r, checkpoint_segments=segments, device=dev, **kw)
    recs = {k: res[k].detach() for k in ("p", "u", "w")}
    g = None
    if grad:
        o = obs[torch.as_tensor(sel, device=obs.device)].to(device=dev, dtype=dtype)
        _misfit(not kw, loss, wl.dt)(o, res["p"]).backward()
        g = v.grad.detach()
    return recs, res["forward_wavefield_p"].detach(), g


def _trace_report(tag, obs, syn, g):
    """Prints the traces whose residual energy is zero (their misfit gradient is NaN, as in the
    reference) and the NaN count of a gradient."""
    e = ((obs.double() - syn.double()) ** 2).sum(dim=1)
    print(f"{tag}: traces {e.numel()}, zero-residual traces {int((e == 0).sum())}, "
          f"min residual energy {float(e.min()):.3e}, NaN in gradient {int(torch.isnan(g).sum())}", flush=True)


def _acoustic_setup(wl):
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.propagator.boundary_condition import bc_pml
    m = wl.models
    damp = torch.tensor(bc_pml(wl.nx, wl.nz, wl.dx, wl.dx, wl.nabc, float(m["vp_init"].max()), False),
                        dtype=torch.float64)
    allshots = np.arange(wl.nshots)
    with torch.no_grad():
        obs, _, _ = _acoustic(ak, DEV, wl, m["vp_true"], m["rho_true"], damp, None, allshots, grad=False)
    return damp, obs["p"], allshots


@pytest.mark.timeout(300)
def test_acoustic_100x200_10_shots_against_the_cpu_oracle():
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.workloads import acoustic_small
    from oracle import acoustic as oa
    wl = acoustic_small()
    m = wl.models
    damp, obs, shots = _acoustic_setup(wl)
    t0 = time.perf_counter()
    h_rec, h_fw, h_g = _acoustic(ak, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, shots, loss="l2")
    torch.cuda.synchronize()
    _trace_report("acoustic 100x200", obs, h_rec["p"], h_g)
    obs_c = obs.cpu()
    o32_rec, o32_fw, o32_g = _acoustic(oa, "cpu", wl, m["vp_init"], m["rho_init"], damp, obs_c, shots, loss="l2")
    o64_rec, o64_fw, o64_g = _acoustic(oa, "cpu", wl, m["vp_init"], m["rho_init"], damp, obs_c, shots,
                                       dtype=torch.float64, loss="l2")
    print(f"acoustic 100x200: {time.perf_counter() - t0:.1f} s", flush=True)
    cfg = "acoustic 100x200, 10 shots, nt 1000"
    for k in ("p", "u", "w"):
        e32, e64 = _rel(h_rec[k].cpu(), o32_rec[k]), _rel(h_rec[k].cpu(), o64_rec[k])
        _log(cfg, "records " + k, hip_vs_fp32=e32, hip_vs_fp64=e64, fp32_vs_fp64=_rel(o32_rec[k], o64_rec[k]))
        assert e32 <= REC_BAR and e64 <= FP64_LOOSE, (k, e32, e64)
    e32, e64 = _rel(h_fw.cpu(), o32_fw), _rel(h_fw.cpu(), o64_fw)
    _log(cfg, "forward_wavefield_p", hip_vs_fp32=e32, hip_vs_fp64=e64, fp32_vs_fp64=_rel(o32_fw, o64_fw))
    assert e32 <= REC_BAR, e32
    e32, e64 = _rel(h_g.cpu(), o32_g), _rel(h_g.cpu(), o64_g)
    _log(cfg, "vp gradient", hip_vs_fp32=e32, hip_vs_fp64=e64, fp32_vs_fp64=_rel(o32_g, o64_g))
    assert e32 <= GRAD_BAR and e64 <= FP64_LOOSE, (e32, e64)


@pytest.mark.timeout(600)
def test_acoustic_384x1152_64_shots_nt4000_against_the_oracle_on_the_gpu():
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.workloads import acoustic_large
    from oracle import acoustic as oa
    wl = acoustic_large()
    m = wl.models
    cfg = "acoustic 384x1152, 64 shots, nt 4000"
    damp, obs, shots = _acoustic_setup(wl)
    h_rec, h_fw, h_g = _acoustic(ak, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, shots)
    assert ak.checkpoint_interval(wl.nt) == 64
    torch.cuda.synchronize()
    _trace_report("acoustic 384x1152", obs, h_rec["p"], h_g)
    # the plain autograd path of the reference algorithm on the GPU, 64 checkpoint segments
    t0 = time.perf_counter()
    o_rec, o_fw, o_g = _acoustic(oa, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, shots, segments=64)
    torch.cuda.synchronize()
    print(f"acoustic 384x1152: fp32 oracle {time.perf_counter() - t0:.1f} s", flush=True)
    errs = {k: _rel(h_rec[k], o_rec[k]) for k in ("p", "u", "w")}
    e_fw, e_g = _rel(h_fw, o_fw), _rel(h_g, o_g)
    # every eighth shot: the oracle in FP64 (and fp32) against the HIP path on the same shots
    sub = shots[::8]
    s_rec, _, s_g = _acoustic(ak, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, sub)
    del o_rec, o_fw
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    q_rec, _, q_g = _acoustic(oa, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, sub, dtype=torch.float64,
                              segments=64)
    torch.cuda.synchronize()
    print(f"acoustic 384x1152: fp64 oracle on {len(sub)} shots {time.perf_counter() - t0:.1f} s", flush=True)
    r_rec, _, r_g = _acoustic(oa, DEV, wl, m["vp_init"], m["rho_init"], damp, obs, sub, segments=64)
    for k in ("p", "u", "w"):
        _log(cfg, "records " + k, hip_vs_fp32=errs[k], hip_vs_fp64=_rel(s_rec[k], q_rec[k]),
             fp32_vs_fp64=_rel(r_rec[k], q_rec[k]), fp64_sample=f"{len(sub)} of 64 shots")
    _log(cfg, "forward_wavefield_p", hip_vs_fp32=e_fw)
    e64 = _rel(s_g, q_g)
    _log(cfg, "vp gradient", hip_vs_fp32=e_g, hip_vs_fp64=e64, fp32_vs_fp64=_rel(r_g, q_g),
         fp64_sample=f"{len(sub)} of 64 shots", hip_subset_vs_fp32_subset=_rel(s_g, r_g))
    del q_rec, r_rec, s_rec
    torch.cuda.empty_cache()
    # central difference of the misfit along a smooth direction (at most 2 m/s), HIP forward only
    d = _direction(11, (wl.nz, wl.nx)) * 2.0
    vals = []
    with torch.no_grad():
        for sgn in (1.0, -1.0):
            rec, _, _ = _acoustic(ak, DEV, wl, (m["vp_init"] + sgn * d).astype(np.float32), m["rho_init"], damp,
                                  obs, shots, grad=False)
            vals.append(_lsq_f64(obs, rec["p"], wl.dt))
    fd = (vals[0] - vals[1]) / 2.0
    dd = float((h_g.double() * torch.tensor(d, device=DEV)).sum())
    e_fd = abs(fd - dd) / abs(dd)
    _log(cfg, "vp gradient . d vs central difference", fd_rel=e_fd)
    for k in ("p", "u", "w"):
        assert errs[k] <= REC_BAR, (k, errs[k])
    assert e_fw <= REC_BAR, e_fw
    assert e_g <= GRAD_BAR and e64 <= FP64_LOOSE, (e_g, e64)
    assert e_fd <= FD_BAR, (fd, dd)


# ----------------------------------------------------------------------------- elastic

LEAVES = ("c11", "c13", "c33", "c55", "bx", "bz")
ELASTIC_NAMES = ("txx", "tzz", "txz", "vx", "vz")


def _elastic_leaves(wl, smooth):
    from oracle import model as om
    m = wl.models
    f = (lambda a: gaussian_filter(a, 10, mode="reflect").astype(np.float32)) if smooth else (lambda a: a)
    t = lambda k: torch.tensor(f(m[k])) if k in m else torch.zeros((wl.nz, wl.nx))
    CC, bx, bz, _, _ = om.elastic_cc(t("vp"), t("vs"), t("rho"), t("eps"), t("delta"),
                                     torch.zeros((wl.nz, wl.nx)), wl.nx, wl.nz)
    return {"c11": CC[0], "c13": CC[2], "c33": CC[11], "c55": CC[18], "bx": bx, "bz": bz}


def _elastic(mod, dev, wl, leaves, shots, obs=None, dtype=torch.float32, segments=1, grad=True, delta=None):
    """forward_kernel of `mod` with explosive moment-tensor shots; L2 misfit of vx and vz
    against obs; returns (records, gradients of the six coefficient fields or None).
    delta: optional perturbation added to the leaves (finite differences)."""
    from adfwi_amd.propagator.boundary_condition import bc_pml_xz
    kw = {} if mod.__name__.startswith("adfwi_amd") else {"dtype": dtype}
    L = {}
    for k in LEAVES:
        x = leaves[k].to(device=dev, dtype=dtype)
        if delta is not None:
            x = x + delta[k].to(device=dev, dtype=dtype)
        L[k] = x.clone().requires_grad_(grad)
    nz, nx = wl.nz, wl.nx
    z = torch.zeros((nz, nx), dtype=dtype, device=dev)
    CC = [L["c11"], z, L["c13"], z, z, z, z, z, z, z, z, L["c33"], z, z, z, z, z, z, L["c55"], z, z]
    vmax = float(wl.models["vp"].max())
    bcx, bcz = bc_pml_xz(nx, nz, wl.dx, wl.dx, wl.nabc, vmax, wl.free_surface)
    sel = np.asarray(shots)
    S = len(sel)
    res = mod.forward_kernel(nx, nz, wl.dx, wl.dx, wl.nt, wl.dt, wl.nabc, wl.free_surface,
                             torch.tensor(wl.src_x[sel], device=dev), torch.tensor(wl.src_z[sel], device=dev), S,
                             torch.tensor(1e6 * wl.wavelets()[sel], dtype=dtype, device=dev),
                             torch.eye(3, dtype=dtype).repeat(S, 1, 1).to(dev),
                             torch.tensor(wl.rcv_x, device=dev), torch.tensor(wl.rcv_z, device=dev), len(wl.rcv_x),
                             "PML", torch.tensor(bcx, dtype=torch.float32, device=dev),
                             torch.tensor(bcz, dtype=torch.float32, device=dev),
                             None, z, z, L["bx"], L["bz"], CC, n_segments=segments, device=dev, **kw)
    recs = {k: res[k].detach() for k in ELASTIC_NAMES}
    if not grad:
        return recs, None
    f = _misfit(not kw, "lsq", wl.dt)
    sum(f(obs[k].to(device=dev, dtype=dtype), res[k]) for k in ("vx", "vz")).backward()
    return recs, {k: L[k].grad.detach() for k in LEAVES}


def _elastic_config(wl, cfg, shots):
    from adfwi_amd.propagator import elastic_kernels as ek
    from oracle import elastic as oe
    true, init = _elastic_leaves(wl, False), _elastic_leaves(wl, True)
    with torch.no_grad():
        obs, _ = _elastic(ek, DEV, wl, true, shots, grad=False)
    h_rec, h_g = _elastic(ek, DEV, wl, init, shots, obs)
    torch.cuda.synchronize()
    _trace_report(cfg + ", vz", obs["vz"], h_rec["vz"], h_g["c33"])
    res = {}
    for dtype in (torch.float32, torch.float64):
        t0 = time.perf_counter()
        res[dtype] = _elastic(oe, DEV, wl, init, shots, obs, dtype=dtype, segments=64)
        torch.cuda.synchronize()
        print(f"{cfg}: {dtype} oracle {time.perf_counter() - t0:.1f} s", flush=True)
    (o_rec, o_g), (q_rec, q_g) = res[torch.float32], res[torch.float64]
    fails = []
    for k in ELASTIC_NAMES:
        e32, e64 = _rel(h_rec[k], o_rec[k]), _rel(h_rec[k], q_rec[k])
        _log(cfg, "records " + k, hip_vs_fp32=e32, hip_vs_fp64=e64, fp32_vs_fp64=_rel(o_rec[k], q_rec[k]))
        if not (e32 <= REC_BAR and e64 <= FP64_LOOSE):
            fails.append(("records", k, e32, e64))
    for k in LEAVES:
        e32, e64 = _rel(h_g[k], o_g[k]), _rel(h_g[k], q_g[k])
        _log(cfg, k + " gradient", hip_vs_fp32=e32, hip_vs_fp64=e64, fp32_vs_fp64=_rel(o_g[k], q_g[k]))
        if not (e32 <= GRAD_BAR and e64 <= FP64_LOOSE):
            fails.append(("gradient", k, e32, e64))
    del res, o_rec, q_rec
    torch.cuda.empty_cache()
    # central difference along a smooth direction: each field scaled by 1e-3 of itself
    rng_dir = {k: torch.tensor(_direction(21 + i, tuple(init[k].shape), 10.0), dtype=torch.float32)
               for i, k in enumerate(LEAVES)}
    delta = {k: 1e-3 * init[k] * rng_dir[k] for k in LEAVES}
    vals = []
    with torch.no_grad():
        for sgn in (1.0, -1.0):
            rec, _ = _elastic(ek, DEV, wl, init, shots, grad=False, delta={k: sgn * v for k, v in delta.items()})
            vals.append(sum(_lsq_f64(obs[k], rec[k], wl.dt) for k in ("vx", "vz")))
    fd = (vals[0] - vals[1]) / 2.0
    dd = sum(float((h_g[k].double() * delta[k].to(DEV).double()).sum()) for k in LEAVES)
    e_fd = abs(fd - dd) / abs(dd)
    _log(cfg, "coefficient gradients . d vs central difference", fd_rel=e_fd)
    assert not fails, fails
    assert e_fd <= FD_BAR, (fd, dd)


@pytest.mark.timeout(600)
def test_isotropic_elastic_200x600_nt4000_against_the_oracle_on_the_gpu():
    from adfwi_amd.workloads import elastic_iso
    wl = elastic_iso()
    _elastic_config(wl, "isotropic elastic 200x600, free surface, nt 4000, 4 explosive shots",
                    np.array([0, 10, 21, 31]))


@pytest.mark.timeout(600)
def test_vti_elastic_320x960_nt5000_against_the_oracle_on_the_gpu():
    from adfwi_amd.workloads import elastic_vti
    wl = elastic_vti()
    _elastic_config(wl, "VTI elastic 320x960, free surface, nt 5000, 2 explosive shots", np.array([5, 40]))


# ----------------------------------------------------------------------------- serving

@pytest.mark.timeout(600)
def test_serving_stream_sample_alone_equals_mix():
    """The first 20 requests of the seeded 2000-request stream at its own ranges: the serving
    loop (continuous batching, footprint admission, same-model packing) against each request
    run alone on the default stream.  Modelling records bit-identical; gradients and losses
    within 1e-6 (rel-L2)."""
    import copy
    from adfwi_amd.serving import ServingLoop, execute, make_stream
    reqs = make_stream(20, 1.0, seed=4)
    assert {r.kind for r in reqs} == {"acoustic", "elastic"} and {r.mode for r in reqs} == {"forward", "gradient"}
    for r in reqs:
        r.keep = True
    t0 = time.perf_counter()
    rep = ServingLoop(DEV, max_slots=4).run(copy.deepcopy(reqs), realtime=False)
    print(f"serving sample: mix {time.perf_counter() - t0:.1f} s, max concurrent {rep.max_concurrent}", flush=True)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    assert rep.max_concurrent >= 2
    worst = {"forward": 0.0, "gradient": 0.0}
    for o in rep.outcomes:
        r = reqs[o.rid]
        alone, _ = execute(r, DEV)
        torch.cuda.synchronize()
        assert o.result.keys() == alone.keys(), (o.rid, o.result.keys(), alone.keys())
        for k in alone:
            x, y = o.result[k], alone[k]
            e = _rel(x, y)
            worst[r.mode] = max(worst[r.mode], e)
            if r.mode == "forward":
                assert torch.equal(x, y), (o.rid, k, e)
            else:
                assert float(y.abs().max()) > 0 and e <= 1e-6, (o.rid, k, e)
        o.result = None
        del alone
    _log("serving stream, first 20 of 2000 requests (seed 4)", "mix vs alone",
         forward_max_rel=worst["forward"], gradient_max_rel=worst["gradient"])
