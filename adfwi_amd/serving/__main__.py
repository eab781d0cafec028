"""Runs a seeded request stream through the serving loop at several arrival rates and, beside
it, the plain path (one request at a time, first come first served); prints one JSON line.

    python -m adfwi_amd.serving --requests 60 --load 0.5,1,2 --slots 4 [--scale-nt 0.1 --max-shots 8]
                                [--model-pool 6 | --shape-pool 4] [--graph-steps 64] [--no-pack]
                                [--pack-across-models] [--exclusive-s 1.0] [--burst]
                                [--slots 4,2 --one-long-s 2.0 --long-slots-s 2.0]

The plain path is run once without arrivals to measure every request's service time; its
capacity (requests/s) sets the arrival rates (--load x capacity).  Its latency at each rate is
that of a single FIFO server with those service times and the same arrivals.  The scale
options shrink nt, the shot count and the grid of every request so that a bounded run finishes
in minutes; the JSON line states them.  --model-pool N draws every request against one of N
shared models (the shot requests of an inversion), so that requests of one model and mode are
packed into shared launches (--no-pack turns packing off); --shape-pool N draws every request
with one of N shapes but its own model, so that acoustic modelling requests of one shape pack
(and, with --pack-across-models, acoustic gradient requests of one shape); --graph-steps sets the HIP-graph
window of acoustic modelling (0: direct launches).  --exclusive-s S gives every request
estimated to take longer than S seconds alone the device to itself.  --burst submits the whole stream at once
(saturated throughput; the plain FIFO then starts every request at t = 0 too).  Several
serving policies run on the same stream after one plain pass: the loop's defaults when no
policy is named, else every --slots value (a comma list; one slot per request), with --one-long-s S also the loop that runs requests estimated above S seconds one at
a time beside shorter ones (4 slots, each such request charged 0.51 of the budget), and with
--long-slots-s S the loop whose requests estimated above S seconds take two of its four slots.
"""
import argparse
import copy
import json
import sys
import time

import numpy as np


def fifo_latencies(arrivals, service):
    free, out = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        out.append(free - a)
    return np.array(out)


def measure_stream(requests, loads, *, slots=4, seed=4, nz_range=(128, 512), nx_range=(256, 2048),
                   nt_range=(1000, 6000), shots_range=(1, 64), model_pool=0, shape_pool=0, graph_steps=64,
                   pack=True, pack_across_models=False, exclusive_s=None, long_charge=1.0, burst=False,
                   device="cuda:0", log=None, policies=None):
    """Runs the seeded stream through the plain path (every request alone, one after the other:
    the service times and the capacity) and then through the serving loop at each load (arrival
    rate = load x plain capacity); returns the report dictionary.  policies: a list of
    (name, ServingLoop keyword arguments) run one after another at every load (default: one
    loop with `slots`, `exclusive_s` and `long_charge`)."""
    from .loop import ServingLoop, run_alone
    from .requests import make_stream
    log = log or (lambda m: print(f"[serving] {m}", file=sys.stderr, flush=True))
    last = [0.0]

    def progress(what):
        def cb(done, total):   # a line at most every 30 s, and at the end
            now = time.perf_counter()
            if done == total or now - last[0] >= 30.0:
                last[0] = now
                log(f"{what}: {done}/{total} requests done")
        return cb

    def stream(rate):
        return make_stream(requests, rate, seed, nz_range=nz_range, nx_range=nx_range, nt_range=nt_range,
                           shots_range=shots_range, model_pool=model_pool, shape_pool=shape_pool)
    # warm-up outside the measurement: one small request of each kind and mode
    warm = make_stream(4, 0.0, 99, nz_range=(128, 128), nx_range=(256, 256), nt_range=(40, 40),
                       shots_range=(1, 1), acoustic_frac=0.5, gradient_frac=0.5)
    ServingLoop(device, max_slots=2, graph_steps=graph_steps).run(warm, realtime=False)
    base = stream(1.0)
    t0 = time.perf_counter()
    alone = run_alone(copy.deepcopy(base), device, progress=progress("plain path"))
    plain_wall = time.perf_counter() - t0
    service = np.array([o.end - o.start for o in alone])
    capacity = len(base) / service.sum()
    log(f"plain path: {len(base)} requests in {plain_wall:.2f} s, capacity {capacity:.2f}/s")
    out = {"metric": "serving latency (p50/p99, s) and throughput (requests/s)",
           "stream": {"requests": requests, "seed": seed,
                      "mix": "60% acoustic / 40% isotropic elastic, 50% forward / 50% gradient",
                      "nz": list(nz_range), "nx": list(nx_range), "nt": list(nt_range),
                      "shots": list(shots_range), "arrivals": "Poisson", "model_pool": model_pool,
                      "shape_pool": shape_pool},
           "slots": slots, "packing": pack, "pack_across_models": pack_across_models,
           "exclusive_s": exclusive_s, "long_charge": long_charge,
           "graph_steps": graph_steps,
           "plain_capacity_req_per_s": capacity,
           "plain_service_p50_s": float(np.percentile(service, 50)),
           "plain_service_p99_s": float(np.percentile(service, 99)), "runs": []}
    digests = {o.rid: o.digest for o in alone}
    if policies is None:   # the loop's defaults otherwise (long requests take two slots)
        policies = [("default", dict(max_slots=slots, exclusive_s=exclusive_s, long_charge=long_charge))]
    out["policies"] = {name: {k: v for k, v in kw.items()} for name, kw in policies}
    for load, (pname, pkw) in [(ld, pol) for ld in ([0.0] if burst else list(loads)) for pol in policies]:
        rate = load * capacity
        reqs = stream(rate if not burst else 1.0)
        if burst:
            for r in reqs:
                r.arrival = 0.0
        rep = ServingLoop(device, pack=pack, pack_across_models=pack_across_models, graph_steps=graph_steps,
                          **pkw).run(reqs, realtime=not burst, progress=progress(f"{pname} at load {load}"))
        s = rep.summary()
        fifo = fifo_latencies([r.arrival for r in reqs], service)
        same = all(o.digest.keys() == digests[o.rid].keys() and all(
            abs(o.digest[k][1] - digests[o.rid][k][1]) <= 1e-6 * max(abs(digests[o.rid][k][1]), 1e-30)
            for k in o.digest) for o in rep.outcomes if o.error is None)
        run = {"load": "burst" if burst else load, "policy": pname, "rate_req_per_s": None if burst else rate,
               "serving": s, "latency_max_s": float(max(o.latency for o in rep.outcomes)),
               "plain_fifo": {"latency_p50_s": float(np.percentile(fifo, 50)),
                              "latency_p99_s": float(np.percentile(fifo, 99)), "latency_max_s": float(fifo.max()),
                              "throughput_req_per_s": len(reqs) / max(fifo[-1] + reqs[-1].arrival, 1e-9)},
               "results_match_plain": bool(same)}
        out["runs"].append(run)
        log(f"{pname} load {load}: p50 {s['latency_p50_s']:.3f} s, p99 {s['latency_p99_s']:.3f} s "
            f"(plain FIFO p50 {run['plain_fifo']['latency_p50_s']:.3f}, p99 {run['plain_fifo']['latency_p99_s']:.3f})")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=60)
    ap.add_argument("--load", default="0.5,1,2", help="arrival rates as multiples of the plain capacity")
    ap.add_argument("--slots", default=None,
                    help="slots of the serving loop, one slot per request (a comma list: one policy each); "
                         "without it and the other policy options, the loop's defaults")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--scale-nt", type=float, default=0.1)
    ap.add_argument("--max-shots", type=int, default=8)
    ap.add_argument("--max-nz", type=int, default=256)
    ap.add_argument("--max-nx", type=int, default=768)
    ap.add_argument("--model-pool", type=int, default=0)
    ap.add_argument("--shape-pool", type=int, default=0,
                    help="requests take one of N (kind, grid, nt) shapes, each with its own model")
    ap.add_argument("--graph-steps", type=int, default=64)
    ap.add_argument("--no-pack", action="store_true")
    ap.add_argument("--pack-across-models", action="store_true",
                    help="acoustic gradient requests of one grid and nt pack whatever their models")
    ap.add_argument("--exclusive-s", type=float, default=None,
                    help="requests estimated to take longer than this alone run with the device to themselves")
    ap.add_argument("--long-charge", type=float, default=1.0,
                    help="fraction of the budget charged to those requests (0.51: one at a time)")
    ap.add_argument("--one-long-s", type=float, default=None,
                    help="also run the policy that serves requests estimated above this many seconds one at a time")
    ap.add_argument("--long-slots-s", type=float, default=None,
                    help="also run the policy whose requests estimated above this many seconds take two of four slots")
    ap.add_argument("--repeat", type=int, default=1, help="run every policy this many times, one after another")
    ap.add_argument("--burst", action="store_true")
    args = ap.parse_args()
    policies = [(f"slots{n}", dict(max_slots=n, exclusive_s=args.exclusive_s, long_charge=args.long_charge,
                                   long_slots_s=None))
                for n in (int(x) for x in args.slots.split(","))] if args.slots else []
    if args.one_long_s is not None:
        policies.append((f"one_long_above_{args.one_long_s:g}s",
                         dict(max_slots=4, exclusive_s=args.one_long_s, long_charge=0.51)))
    if args.long_slots_s is not None:
        policies.append((f"two_slots_above_{args.long_slots_s:g}s", dict(max_slots=4, long_slots_s=args.long_slots_s)))
    if not policies:       # the loop's defaults (four slots, requests estimated above 0.5 s take two)
        policies = [("default", dict(exclusive_s=args.exclusive_s, long_charge=args.long_charge))]
    policies = [(name if i == 0 else f"{name}#{i + 1}", kw) for name, kw in policies for i in range(args.repeat)]
    nt_range = (max(50, int(1000 * args.scale_nt)), max(60, int(6000 * args.scale_nt)))
    out = measure_stream(args.requests, [float(x) for x in args.load.split(",")], seed=args.seed, policies=policies,
                         nz_range=(128, args.max_nz), nx_range=(256, args.max_nx), nt_range=nt_range,
                         shots_range=(1, args.max_shots), model_pool=args.model_pool, shape_pool=args.shape_pool,
                         graph_steps=args.graph_steps, pack=not args.no_pack,
                         pack_across_models=args.pack_across_models, burst=args.burst)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
