"""Benchmark of the propagator hot path on one MI355X.

A step is one FWI gradient of the whole shot batch: forward modelling of all
shots (records, energy maps, sqrt(nt) wavefield checkpoints), the least-squares
waveform misfit 1/2 dt sum (obs - syn)^2 (value and adjoint source in one
kernel), and the hand-written adjoint (segment recomputation + time-reversed
adjoint + zero-lag imaging) down to vp.grad.  The reference's per-trace-root L2
(ADFWI/fwi/misfit/L2.py:26-29) is not used here: 7712 of the 73728 traces of the
default workload (far offsets of the edge shots) receive no arrival within 4 s,
their residual is exactly zero and that misfit's gradient is NaN everywhere.

Default workload (acoustic, one launch of 64 shots): grid 384 x 1152,
dx 12.5 m, 40-cell PML, dt 1 ms, nt 4000, 8 Hz Ricker, 1152 receivers,
checkpoint interval ceil(sqrt(nt)) = 64.  Synthetic seeded models; the
observed data are modelled from the true model before timing.

Prints one JSON line: metric, value (shot-gradients/s), roofline of the
dominant kernel (HIP-event timed inside the timed region; algorithmic and
counted-traffic fractions), cpu_baseline (the CPU oracle on the 100 x 200 /
10-shot / nt 1000 workload, all steps, median of 3), the plain-autograd path on
the same GPU (the whole workload, ceil(sqrt(nt)) checkpoint segments) and, after
all timing, a serving leg: p50/p99 latency and throughput of a bounded stream
drawn from the 2000-request serving workload's ranges at a Poisson rate of 1x
the plain path's measured capacity, beside that plain path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload acoustic_large]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per padded point, per time step, per shot (fp32 state, imaging traffic)
BYTES_FWD = 24.0          # forward: read + write p, u, w
BYTES_RECOMPUTE = 28.0    # recompute: forward + stencil image store
BYTES_ADJ = 28.0          # adjoint: read + write adjoint p, u, w + stencil image read
BYTES_GRAD = 80.0         # per shot-gradient: 24 + 28 + 28
# elastic: ten split fields
EL_BYTES_FWD = 80.0       # forward / recompute: read + write the ten splits
EL_BYTES_ADJ = 120.0      # adjoint: read + write ten adjoint splits, read the forward state before and after
EL_BYTES_GRAD = 280.0     # per shot-gradient: 80 + 80 + 120


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 counter
    summaries (lines `traffic_bytes_per_launch <kernel> <bytes> <workload>`), or None."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc_traffic*.md"))):
        for line in open(f):
            parts = line.split()
            if (len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel
                    and parts[3] == workload):
                return float(parts[2])
    return None


def misfit(syn, obs, dt):
    """Least-squares waveform misfit 1/2 dt sum (obs - syn)^2, its value and adjoint source from
    one kernel (adfwi_amd.fwi.Misfit_waveform_L2_squared)."""
    from adfwi_amd.fwi import Misfit_waveform_L2_squared
    return Misfit_waveform_L2_squared(dt).forward(obs, syn)


def build_problem(wl, dev):
    from adfwi_amd.model import AcousticModel
    from adfwi_amd.propagator import AcousticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    src = Source(wl.nt, wl.dt, wl.f0)
    wav = wl.wavelets()[0]
    src.add_sources(wl.src_x, wl.src_z, wav)
    rcv = Receiver(wl.nt, wl.dt)
    rcv.add_receivers(wl.rcv_x, wl.rcv_z, "pr")
    survey = Survey(src, rcv)
    true = AcousticModel(0, 0, wl.nx, wl.nz, wl.dx, wl.dx, vp=wl.models["vp_true"], rho=wl.models["rho_true"],
                         nabc=wl.nabc, free_surface=wl.free_surface, device=dev)
    init = AcousticModel(0, 0, wl.nx, wl.nz, wl.dx, wl.dx, vp=wl.models["vp_init"], rho=wl.models["rho_init"],
                         vp_grad=True, nabc=wl.nabc, free_surface=wl.free_surface, device=dev)
    prop = AcousticPropagator(init, survey, device=dev)
    with torch.no_grad():
        obs = prop.forward(model=true)["p"].detach().clone()
    return prop, init, obs


def build_elastic_problem(wl, dev, fd_order=4, abc="PML"):
    """Elastic problem of an elastic workload: vertical-force Ricker shots; the true model gives
    the observed vx / vz, the initial model (each field smoothed, sigma 10 cells) carries the
    gradients; rho is a leaf."""
    from scipy.ndimage import gaussian_filter
    from adfwi_amd.model import AnisotropicElasticModel, IsotropicElasticModel
    from adfwi_amd.propagator import ElasticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    src = Source(wl.nt, wl.dt, wl.f0)
    src.add_sources(wl.src_x, wl.src_z, wl.wavelets()[0], "vf")
    rcv = Receiver(wl.nt, wl.dt)
    rcv.add_receivers(wl.rcv_x, wl.rcv_z, "vz")
    survey = Survey(src, rcv)
    smooth = lambda a: gaussian_filter(a, 10, mode="reflect").astype(np.float32)
    m = wl.models
    common = dict(free_surface=wl.free_surface, nabc=wl.nabc, auto_update_rho=False, abc_type=abc, device=dev)
    if "eps" in m:
        zero = np.zeros_like(m["vp"])
        mk = lambda f, grad: AnisotropicElasticModel(
            0, 0, wl.nx, wl.nz, wl.dx, wl.dx, vp=f(m["vp"]), vs=f(m["vs"]), rho=f(m["rho"]), eps=f(m["eps"]),
            gamma=zero, delta=f(m["delta"]), vp_grad=grad, vs_grad=grad, rho_grad=grad, eps_grad=grad,
            delta_grad=grad, **common)
    else:
        mk = lambda f, grad: IsotropicElasticModel(
            0, 0, wl.nx, wl.nz, wl.dx, wl.dx, vp=f(m["vp"]), vs=f(m["vs"]), rho=f(m["rho"]),
            vp_grad=grad, vs_grad=grad, rho_grad=grad, **common)
    true, init = mk(lambda a: a, False), mk(smooth, True)
    prop = ElasticPropagator(init, survey, device=dev)
    with torch.no_grad():
        res = prop.forward(model=true, fd_order=fd_order)
        obs = {k: res[k].detach().clone() for k in ("vx", "vz")}
    return prop, init, obs


def grad_step(prop, model, obs, dt, fd_order=4):
    for p in model.parameters():
        p.grad = None
    rec = prop.forward(fd_order=fd_order) if fd_order != 4 else prop.forward()
    if isinstance(obs, dict):
        loss = sum(misfit(rec[k], obs[k], dt) for k in obs)
    else:
        loss = misfit(rec["p"], obs, dt)
    loss.backward()
    return loss


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_cores():
    """(threads, host physical cores): the threads the CPU baseline uses are the cores this process
    may run on (its affinity), at most OMP_NUM_THREADS when that is set (a GPU's share of the host
    on a shared box); the host's physical core count is reported beside it."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    phys = set()
    try:
        pid = cid = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                    phys.add((pid, cid))
    except OSError:
        pass
    return n, (len(phys) or None)


def cpu_baseline(repeats=3):
    """The CPU oracle (eager torch restatement of the reference path, fp32, the reference's op
    order, autograd) on the 100 x 200 / 10-shot / nt 1000 workload: forward of all shots, the
    least-squares misfit against data the oracle models from the true model, and the backward to
    vp.grad; one short warm-up, then the median of `repeats` full runs."""
    from adfwi_amd.propagator.boundary_condition import bc_pml
    from adfwi_amd.workloads import acoustic_small
    from oracle import acoustic as oa
    from oracle import misfit as om
    cores, phys = cpu_cores()
    torch.set_num_threads(cores)
    wl = acoustic_small()
    nz, nx, nabc = wl.nz, wl.nx, wl.nabc
    damp = torch.tensor(bc_pml(nx, nz, wl.dx, wl.dx, nabc, float(wl.models["vp_init"].max()), False),
                        dtype=torch.float32)
    wav = torch.tensor(wl.wavelets())
    args = (torch.tensor(wl.src_x), torch.tensor(wl.src_z), wl.nshots)
    rcv = (torch.tensor(wl.rcv_x), torch.tensor(wl.rcv_z), len(wl.rcv_x), damp)

    def forward(vp, rho, nt_run):
        return oa.forward_kernel(nx, nz, wl.dx, wl.dx, nt_run, wl.dt, nabc, False, *args, wav[:, :nt_run], *rcv,
                                 vp, rho, checkpoint_segments=1)["p"]
    with torch.no_grad():
        obs = forward(torch.tensor(wl.models["vp_true"]), torch.tensor(wl.models["rho_true"]), wl.nt)
    rho = torch.tensor(wl.models["rho_init"])

    def one(nt_run):
        v = torch.tensor(wl.models["vp_init"], requires_grad=True)
        om.waveform_l2_squared(obs[:, :nt_run], forward(v, rho, nt_run), wl.dt).backward()
        return v.grad
    one(100)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        one(wl.nt)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    N = (nz + 2 * nabc) * (nx + 2 * nabc)
    return {"value": wl.nshots / t, "unit": "shot-gradients/s", "cores": cores, "kind": "port",
            "host_physical_cores": phys,
            "sample": (f"oracle (eager torch fp32, reference op order, autograd) forward + least-squares misfit + "
                       f"backward, 100x200 grid, 20-cell PML, 10 shots, all 1000 steps, median of {repeats} "
                       f"({', '.join(f'{x:.2f}' for x in times)} s) after a warm-up, {cores} threads "
                       f"(this process's cores, capped at the GPU's host share)"),
            "gpt_updates_per_s": 2 * wl.nshots * wl.nt * N / t / 1e9,
            "cpu": _cpu_name()}


def elastic_cpu_baseline(wl, seconds_budget=24.0, fd_order=4, abc="PML"):
    """The CPU oracle (eager torch fp32, the reference's op order, autograd) on a bounded sample of
    the elastic workload: all shots cut to `nb` shots and a few time steps, scaled per shot-step."""
    from adfwi_amd.propagator.boundary_condition import bc_gerjan, bc_pml_xz, bc_sincos
    from oracle import elastic as oe
    from oracle import model as om
    cores, _ = cpu_cores()
    torch.set_num_threads(cores)
    m = {k: torch.tensor(v) for k, v in wl.models.items()}
    nz, nx = wl.nz, wl.nx
    zero = torch.zeros_like(m["vp"])
    cc0, bx0, bz0, _, _ = om.elastic_cc(m["vp"], m["vs"], m["rho"], m.get("eps", zero), m.get("delta", zero),
                                        zero, nx, nz)
    bcx, bcz = bc_pml_xz(nx, nz, wl.dx, wl.dx, wl.nabc, float(m["vp"].max()), wl.free_surface)
    damp = None
    if abc == "gerjan":
        damp = torch.tensor(bc_gerjan(nx, nz, wl.dx, wl.dx, wl.nabc, 0.0053, wl.free_surface), dtype=torch.float32)
    elif abc == "sincos":
        damp = torch.tensor(bc_sincos(nx, nz, wl.dx, wl.dx, wl.nabc, wl.free_surface), dtype=torch.float32)
    nb = 2

    def one(nt_run):
        CC = list(cc0)
        for i in (0, 2, 11, 18):
            CC[i] = cc0[i].clone().requires_grad_(True)
        bx = bx0.clone().requires_grad_(True)
        bz = bz0.clone().requires_grad_(True)
        z = torch.zeros((nz, nx))
        wav = torch.tensor(wl.wavelets()[:nb, :nt_run])
        res = oe.forward_kernel(nx, nz, wl.dx, wl.dx, nt_run, wl.dt, wl.nabc, wl.free_surface,
                                torch.tensor(wl.src_x[:nb]), torch.tensor(wl.src_z[:nb]), nb, wav,
                                torch.tensor(np.tile(np.eye(3, dtype=np.float32), (nb, 1, 1))),
                                torch.tensor(wl.rcv_x), torch.tensor(wl.rcv_z), len(wl.rcv_x), abc,
                                torch.tensor(bcx, dtype=torch.float32), torch.tensor(bcz, dtype=torch.float32),
                                damp, z, z, bx, bz, CC, fd_order=fd_order)
        (res["vx"].square().sum() + res["vz"].square().sum()).backward()
    t0 = time.perf_counter()
    one(10)
    per_step = (time.perf_counter() - t0) / 10
    nt_run = int(max(10, min(wl.nt, seconds_budget / 3 / max(per_step, 1e-9))))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        one(nt_run)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times)) * (wl.nt / nt_run) * (wl.nshots / nb)
    return {"value": wl.nshots / t, "unit": "shot-gradients/s", "cores": cores, "kind": "port",
            "sample": (f"oracle (eager torch fp32, reference op order, autograd) forward+backward, full "
                       f"{nz}x{nx} grid, fd_order {fd_order}, {abc}, {nb} of {wl.nshots} shots, {nt_run} of {wl.nt} steps, "
                       f"median of 3, "
                       f"scaled per shot and step; explosive moment-tensor shots (the reference has no "
                       f"vertical-force source; the stepping cost is the same)"),
            "cpu": _cpu_name()}


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def plain_gpu_baseline(wl, dev, obs):
    """The oracle (the reference's plain autograd path: eager torch, TorchScript steps, reentrant
    checkpointing) on the GPU for the whole benchmark workload: all shots and steps, ceil(sqrt(nt))
    checkpoint segments, the same least-squares misfit against the same observed data, backward to
    vp.grad; timed once after a short warm-up."""
    from adfwi_amd.propagator.boundary_condition import bc_pml
    from oracle import acoustic as oa
    from oracle import misfit as om
    nz, nx, nabc = wl.nz, wl.nx, wl.nabc
    damp = torch.tensor(bc_pml(nx, nz, wl.dx, wl.dx, nabc, float(wl.models["vp_init"].max()), False),
                        dtype=torch.float32, device=dev)
    wav = torch.tensor(wl.wavelets(), device=dev)
    rho = torch.tensor(wl.models["rho_init"], device=dev)
    segs = int(math.ceil(math.sqrt(wl.nt)))

    def one(nt_run, nseg):
        v = torch.tensor(wl.models["vp_init"], device=dev, requires_grad=True)
        res = oa.forward_kernel(nx, nz, wl.dx, wl.dx, nt_run, wl.dt, nabc, False,
                                torch.tensor(wl.src_x, device=dev), torch.tensor(wl.src_z, device=dev),
                                wl.nshots, wav[:, :nt_run], torch.tensor(wl.rcv_x, device=dev),
                                torch.tensor(wl.rcv_z, device=dev), len(wl.rcv_x), damp, v, rho,
                                checkpoint_segments=nseg, device=dev)
        om.waveform_l2_squared(obs[:, :nt_run], res["p"], wl.dt).backward()
    one(2 * (wl.nt // segs), 2)     # warm-up: TorchScript profiles the step function
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    one(wl.nt, segs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return {"value": wl.nshots / t, "unit": "shot-gradients/s", "seconds": t,
            "sample": f"oracle on the GPU, the whole workload: {wl.nshots} shots, all {wl.nt} steps, "
                      f"{segs} checkpoint segments, least-squares misfit, timed once after a warm-up"}


def serving_leg(requests=32, load=1.0):
    """The serving loop on a bounded stream drawn from the 2000-request serving workload's own
    ranges (seed 4), at a Poisson arrival rate of `load` x the plain path's measured capacity,
    beside the plain path (every request alone, first come first served).  It runs in a process
    of its own (`python -m adfwi_amd.serving`, the loop's default policy), as a serving
    deployment would, and serves the stream twice: a process's first pass pays one-time costs
    that a running server has paid (its p50 measured 2-3x the later passes' at the same
    throughput), so the second pass is the figure and the first is reported beside it."""
    import subprocess
    cmd = [sys.executable, "-m", "adfwi_amd.serving", "--requests", str(requests), "--load", str(load),
           "--scale-nt", "1", "--max-shots", "64", "--max-nz", "512", "--max-nx", "2048", "--repeat", "2"]
    res = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, text=True, check=True)   # progress on stderr
    rep = json.loads(res.stdout.strip().splitlines()[-1])
    first, run = rep["runs"][0], rep["runs"][1]
    s = run["serving"]
    return {"command": "python -m adfwi_amd.serving " + " ".join(cmd[3:]), "stream": rep["stream"],
            "policy": run["policy"], "rate_req_per_s": run["rate_req_per_s"],
            "first_pass": {"latency_p50_s": first["serving"]["latency_p50_s"],
                           "latency_p99_s": first["serving"]["latency_p99_s"],
                           "throughput_req_per_s": first["serving"]["throughput_req_per_s"]},
            "load_x_plain_capacity": load, "latency_p50_s": s["latency_p50_s"], "latency_p99_s": s["latency_p99_s"],
            "throughput_req_per_s": s["throughput_req_per_s"], "completed": s["requests"] - s["failed"],
            "max_concurrent": s.get("max_concurrent"),
            "plain_fifo": run["plain_fifo"], "plain_capacity_req_per_s": rep["plain_capacity_req_per_s"],
            "results_match_plain": run["results_match_plain"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="acoustic_large")
    ap.add_argument("--nt", type=int, default=0, help="override nt (profiling runs only)")
    ap.add_argument("--fd-order", type=int, default=4, choices=(4, 6, 8, 10),
                    help="elastic workloads: FD order (6-10: the three-term schemes)")
    ap.add_argument("--abc", default="PML", choices=("PML", "gerjan", "sincos"),
                    help="elastic workloads: absorbing boundary (gerjan / sincos: the damping scheme)")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=32,
                    help="requests of the serving leg (0: none): a bounded stream from the serving workload's ranges")
    ap.add_argument("--serving-load", type=float, default=1.0,
                    help="serving leg arrival rate as a multiple of the plain path's measured capacity")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many MiB (PMC counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); shots shard across GPUs with "
              "one gradient all-reduce per iteration, which this single-GPU build does not launch",
              file=sys.stderr)
        sys.exit(2)
    dev = "cuda:0"
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe}))
        return

    from adfwi_amd import workloads
    from adfwi_amd.propagator import acoustic_kernels as ak
    kw = {"nt": args.nt} if args.nt else {}
    wl = getattr(workloads, args.workload)(**kw)
    elastic = wl.name.startswith("elastic")
    fd = args.fd_order if elastic else 4
    if fd != 4 and not elastic:
        sys.exit("--fd-order applies to the elastic workloads")
    log(f"workload {wl.name}: grid {wl.nz}x{wl.nx}, nt {wl.nt}, {wl.nshots} shots, fd_order {fd}")
    if args.abc != "PML" and not elastic:
        sys.exit("--abc applies to the elastic workloads")
    prop, model, obs = build_elastic_problem(wl, dev, fd, args.abc) if elastic else build_problem(wl, dev)
    log("observed data modelled")
    for i in range(args.warmup):
        grad_step(prop, model, obs, wl.dt, fd)
        torch.cuda.synchronize()
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    ak.TIMELINE.records.clear()
    ak.TIMELINE.enabled = True
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss = grad_step(prop, model, obs, wl.dt, fd)
        log(f"timed step {i} issued")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ak.TIMELINE.enabled = False
    tl = ak.TIMELINE.summary()
    Z, X = wl.padded_elastic_order(fd) if elastic else wl.padded
    N = Z * X
    S = wl.nshots
    value = S * args.steps / elapsed
    # dominant kernel by total time; its algorithmic bytes per launch
    kern = {}
    for name, d in tl.items():
        steps_by = d["steps_by_call"]
        if name in ("ac_fwd_step", "ac_fwd_pair"):
            byts = (steps_by.get("forward", 0) * BYTES_FWD + (steps_by.get("recompute", 0)
                    + steps_by.get("forward+images", 0)) * BYTES_RECOMPUTE) * N * S
        elif name in ("ac_adj_step", "ac_adj_pair"):
            byts = d["steps"] * BYTES_ADJ * N * S
        elif name in ("el_fwd_step", "el3_fwd_step"):
            byts = d["steps"] * EL_BYTES_FWD * N * S
        elif name in ("el_abl_fwd", "el3_abl_fwd"):
            byts = d["steps"] * EL_BYTES_FWD / 2 * N * S
        elif name in ("el_abl_adj", "el3_abl_adj"):
            byts = d["steps"] * EL_BYTES_ADJ / 2 * N * S
        else:
            byts = d["steps"] * EL_BYTES_ADJ * N * S
        avg_ms = d["ms"] / max(d["launches"], 1)
        kern[name] = {"total_ms": d["ms"], "launches": d["launches"], "steps": d["steps"],
                      "steps_by_call": d["steps_by_call"], "avg_launch_ms": avg_ms,
                      "achieved_gbs": byts / (d["ms"] / 1e3) / 1e9 if d["ms"] else 0.0,
                      "bytes_per_launch": byts / max(d["launches"], 1), "bytes_total": byts}
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    ach = kern[dom]["achieved_gbs"]
    traffic = profiled_traffic(dom, wl.name)
    # bytes of the steps actually executed (resident segments skip part of the recompute) against
    # the specified algorithm's per-pt-step figure
    executed = sum(k["bytes_total"] for k in kern.values())
    spec_bytes = ((EL_BYTES_GRAD / (2 if args.abc != "PML" else 1)) if elastic else BYTES_GRAD) * S * N * wl.nt
    out = {
        "metric": "shot-gradients/s",
        "value": value,
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded models; observed data modelled from the true model)",
        "config": {"workload": wl.name, "grid": [wl.nz, wl.nx], "padded_grid": [Z, X], "nabc": wl.nabc,
                   "nt": wl.nt, "shots_per_launch": S, "receivers": len(wl.rcv_x), "dt": wl.dt, "dx": wl.dx,
                   "checkpoint_interval": ak.checkpoint_interval(wl.nt), "parallelism": "single GPU, shot batch"},
        "gpt_updates_per_s": 2 * S * N * wl.nt * args.steps / elapsed / 1e9,
        # the executed steps' algorithmic bytes over the whole step time, against 8 TB/s
        "grad_bytes_roofline_frac": executed / elapsed / 1e9 / HBM_PEAK_GBS,
        "executed_bytes_per_shot_gradient": executed / (S * args.steps),
        # the specified algorithm's bytes (every segment recomputed) over the same time
        "spec_grad_bytes_roofline_frac": spec_bytes * args.steps / elapsed / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     # counted HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE) over the launch time
                     "traffic_gbs": traffic / (kern[dom]["avg_launch_ms"] / 1e3) / 1e9 if traffic else None,
                     "traffic_frac": (traffic / (kern[dom]["avg_launch_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS
                                      if traffic else None),
                     "bytes_per_launch": kern[dom]["bytes_per_launch"],
                     "avg_launch_ms": kern[dom]["avg_launch_ms"]},
        "kernels": kern,
        "loss": float(loss.detach()),
    }
    log(f"timed region {elapsed:.2f} s, {value:.2f} shot-gradients/s")
    if elastic:
        out["config"]["fd_order"] = fd
        out["config"]["abc_type"] = args.abc
        out["config"]["free_surface"] = wl.free_surface
        out["config"]["components"] = ["vx", "vz"]
    for p in model.parameters():
        p.grad = None
    del prop, model
    torch.cuda.empty_cache()
    if not args.no_baselines and not elastic:
        out["plain_autograd_gpu"] = plain_gpu_baseline(wl, dev, obs)
        log(f"plain autograd GPU baseline done ({out['plain_autograd_gpu']['seconds']:.1f} s)")
        out["speedup_vs_plain_autograd_gpu"] = value / out["plain_autograd_gpu"]["value"]
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline()
        log("CPU baseline done")
    elif not args.no_baselines:
        out["cpu_baseline"] = elastic_cpu_baseline(wl, fd_order=fd, abc=args.abc)
        log("CPU baseline done")
    else:
        out["cpu_baseline"] = None
    del obs
    torch.cuda.empty_cache()
    if args.serving_requests > 0:
        out["serving"] = serving_leg(args.serving_requests, args.serving_load)
        log(f"serving leg done: p50 {out['serving']['latency_p50_s']:.2f} s, "
            f"p99 {out['serving']['latency_p99_s']:.2f} s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
