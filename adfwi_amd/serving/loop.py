"""Continuous-batching serving loop on one GPU.

Requests arrive over time (their `arrival`, seconds after the start).  The loop submits each
arrived request to the Scheduler, which admits it as soon as a slot and its HBM footprint are
free; every admitted request runs on a worker thread with its own HIP stream (default
priority), so requests that are on the GPU together overlap on the device: the library's step
kernels of several requests execute concurrently and the host-side launch loops (which release
the GIL inside the C library) run in parallel.  A request is done when its stream has drained;
its latency is completion time - arrival time.  Model fields, wavefields, checkpoints and
records stay in HBM for the request's lifetime and are freed when it completes.

With packing on (the default), requests against one shared model (`Request.model`) and of one
mode are admitted together and share one sequence of launches (`execute_pack`): their shots
run side by side in the same kernels, and each member gets its own records or gradient.
Acoustic modelling replays HIP graphs of `graph_steps` steps (default 64; 0 = direct
launches): a worker thread keeps the captured graphs of the shapes it has run, so later
requests of the same shape replay them.
"""
import queue
import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .requests import Request, execute, execute_pack, footprint_bytes, pack_key, service_estimate_s
from .scheduler import Scheduler


@dataclass
class Outcome:
    rid: int
    arrival: float
    start: float
    end: float
    digest: dict
    result: Optional[dict] = None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.end - self.arrival


@dataclass
class StreamReport:
    outcomes: List[Outcome]
    rejected: list
    wall_s: float
    peak_admitted_bytes: int
    max_concurrent: int
    packs: List[List[int]] = field(default_factory=list)   # request ids of every admitted pack

    def summary(self) -> dict:
        ok = [o for o in self.outcomes if o.error is None]
        lat = np.array([o.latency for o in ok]) if ok else np.zeros(1)
        return {"requests": len(self.outcomes), "failed": len(self.outcomes) - len(ok),
                "rejected": len(self.rejected), "wall_s": self.wall_s,
                "throughput_req_per_s": len(ok) / self.wall_s if self.wall_s else 0.0,
                "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
                "latency_mean_s": float(lat.mean()), "max_concurrent": self.max_concurrent,
                "peak_admitted_gib": self.peak_admitted_bytes / 2**30,
                "packs": len(self.packs), "packed_requests": sum(len(p) for p in self.packs if len(p) > 1)}


class ServingLoop:
    """Serves a list of requests on `device` with at most `max_slots` in flight and at most
    `budget_bytes` of admitted footprint (default: 85 % of the free device memory).
    `pack_across_models`: acoustic gradient requests of one grid and nt pack whatever their
    models (off by default: on bursts of small requests the fewer, longer packs balance worse
    over the slots than the requests alone).  `exclusive_s`: a request estimated to take longer
    than this alone (`service_estimate_s`) is charged `long_charge` of the whole budget (at least
    its footprint): 1.0 starts it on an idle device and runs it alone; just over 0.5 keeps two
    long requests from sharing the device while short ones still run beside one (None: every
    request is charged its footprint).  `long_slots_s`: a request estimated to take longer than
    this alone takes two slots instead of one, so that at most max_slots / 2 such requests share
    the device (each is slowed less) while short requests still fill every slot (None: one slot
    each).  The default, 0.5 s, is about the median full-size request: on full-size streams at
    the plain capacity it lowered p99 by 7 % against one slot per request at the same
    throughput, and on streams of small requests nothing changes."""

    def __init__(self, device="cuda:0", max_slots: int = 4, budget_bytes: Optional[int] = None,
                 footprint: Callable[[Request], int] = footprint_bytes, max_bypass: int = 4,
                 pack: bool = True, max_pack_shots: int = 64, graph_steps: int = 64,
                 pack_across_models: bool = False, exclusive_s: Optional[float] = None,
                 long_charge: float = 1.0, long_slots_s: Optional[float] = 0.5):
        import torch
        from .. import _native as nat
        nat.require_gpu(device)
        nat.load()
        self.device = torch.device(device)
        if budget_bytes is None:
            # free device memory, including what torch's allocator holds cached but unused (an
            # earlier loop's buffers): requests allocate through that cache
            free, _ = torch.cuda.mem_get_info(self.device)
            free += torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            budget_bytes = int(0.85 * free)
            if graph_steps:       # the graph contexts each worker thread keeps
                from ..propagator import graphs
                budget_bytes -= max_slots * graphs.MAX_CONTEXTS * graphs.MAX_CONTEXT_BYTES
        key = (lambda r: pack_key(r, True)) if pack_across_models else pack_key
        self.scheduler = Scheduler(budget_bytes, max_slots, max_bypass, key if pack else None, max_pack_shots)
        self.footprint = footprint
        if not 0.0 < long_charge <= 1.0:
            raise ValueError(f"long_charge must be in (0, 1] (got {long_charge})")
        self.exclusive_s = exclusive_s
        self.long_charge = float(long_charge)
        self.long_slots_s = long_slots_s
        self.max_slots = max_slots
        self.graph_steps = int(graph_steps)
        self._local = threading.local()

    def _stream(self):
        import torch
        s = getattr(self._local, "stream", None)
        if s is None:
            s = torch.cuda.Stream(device=self.device)   # default priority
            self._local.stream = s
        return s

    def _run_pack(self, reqs: List[Request], t0: float) -> List[Outcome]:
        import torch
        start = time.perf_counter() - t0
        stream = self._stream()
        try:
            with torch.cuda.device(self.device), torch.cuda.stream(stream):
                res = execute_pack(reqs, self.device, self.graph_steps)
                stream.synchronize()
            end = time.perf_counter() - t0
            return [Outcome(r.rid, r.arrival, start, end, dg, out if r.keep else None)
                    for r, (out, dg) in zip(reqs, res)]
        except Exception as exc:  # noqa: BLE001 -- reported per request
            stream.synchronize()
            end = time.perf_counter() - t0
            return [Outcome(r.rid, r.arrival, start, end, {}, None, repr(exc)) for r in reqs]

    def run(self, requests: List[Request], realtime: bool = True,
            progress: Optional[Callable[[int, int], None]] = None) -> StreamReport:
        """Serves the requests.  realtime=False submits every request at t = 0 (arrival times
        are then taken as 0 for the latency).  progress(done, total) is called after every
        completed pack."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        outcomes: List[Outcome] = []
        futures = {}
        max_conc = 0
        packs = []
        t0 = time.perf_counter()
        i = 0
        with ThreadPoolExecutor(max_workers=self.max_slots) as pool:
            while i < len(pending) or futures or self.scheduler.waiting:
                now = time.perf_counter() - t0
                while i < len(pending) and pending[i].arrival <= now:
                    r = pending[i]
                    nbytes = self.footprint(r)
                    if self.exclusive_s is not None and service_estimate_s(r) > self.exclusive_s:
                        # an oversized footprint stays as it is, so that it is rejected
                        nbytes = max(nbytes, int(self.long_charge * self.scheduler.budget))
                    long = self.long_slots_s is not None and service_estimate_s(r) > self.long_slots_s
                    weight = min(2, self.max_slots) if long else 1
                    if not self.scheduler.submit(r, nbytes, weight):
                        outcomes.append(Outcome(r.rid, r.arrival, now, now, {}, None,
                                                self.scheduler.rejected[-1][1]))
                    i += 1
                for pack in self.scheduler.admit_packs():
                    futures[pool.submit(self._run_pack, pack, t0)] = pack
                    packs.append([r.rid for r in pack])
                max_conc = max(max_conc, len(futures))
                timeout = None
                if i < len(pending):
                    timeout = max(0.0, pending[i].arrival - (time.perf_counter() - t0))
                if futures:
                    done, _ = wait(list(futures), timeout=timeout, return_when=FIRST_COMPLETED)
                    for f in done:
                        for r in futures.pop(f):
                            self.scheduler.finish(r.rid)
                        outcomes.extend(f.result())
                        if progress is not None:
                            progress(len(outcomes), len(pending))
                elif timeout is not None:
                    time.sleep(timeout)
                elif self.scheduler.waiting:
                    raise RuntimeError("requests wait but nothing runs: footprint accounting is inconsistent")
        wall = time.perf_counter() - t0
        outcomes.sort(key=lambda o: o.rid)
        return StreamReport(outcomes, list(self.scheduler.rejected), wall, self.scheduler.peak_bytes, max_conc,
                            packs)


def run_alone(requests: List[Request], device="cuda:0",
              progress: Optional[Callable[[int, int], None]] = None) -> List[Outcome]:
    """The plain path: each request by itself, one after the other, on the default stream;
    progress(done, total) after every request."""
    import torch
    out = []
    t0 = time.perf_counter()
    for r in requests:
        s = time.perf_counter() - t0
        res, digest = execute(r, device)
        torch.cuda.synchronize()
        out.append(Outcome(r.rid, r.arrival, s, time.perf_counter() - t0, digest, res if r.keep else None))
        if progress is not None:
            progress(len(out), len(requests))
    return out
