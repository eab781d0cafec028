"""Runs a seeded request stream through the serving loop and, for comparison, the plain path
(one request at a time); prints one JSON line.

    python -m adfwi_amd.serving --requests 40 --rate 2.0 --slots 4 [--scale-nt 0.1 --max-shots 8]

The scale options shrink nt, the shot count and the grid of every request of the stream so
that a bounded run finishes in minutes; the JSON line states them.
"""
import argparse
import json
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=2.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--scale-nt", type=float, default=0.1)
    ap.add_argument("--max-shots", type=int, default=8)
    ap.add_argument("--max-nz", type=int, default=256)
    ap.add_argument("--max-nx", type=int, default=768)
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    from .loop import ServingLoop, run_alone
    from .requests import make_stream
    reqs = make_stream(args.requests, args.rate, args.seed, nz_range=(128, args.max_nz),
                       nx_range=(256, args.max_nx), nt_range=(max(50, int(1000 * args.scale_nt)),
                                                                 max(60, int(6000 * args.scale_nt))),
                       shots_range=(1, args.max_shots))
    loop = ServingLoop("cuda:0", max_slots=args.slots)
    # warm-up: one small request of each kind/mode outside the measurement
    warm = make_stream(4, 0.0, 99, nz_range=(128, 128), nx_range=(256, 256), nt_range=(40, 40),
                       shots_range=(1, 1), acoustic_frac=0.5, gradient_frac=0.5)
    ServingLoop("cuda:0", max_slots=2).run(warm, realtime=False)
    print(f"[serving] {len(reqs)} requests, rate {args.rate}/s, {args.slots} slots", file=sys.stderr, flush=True)
    rep = loop.run(reqs, realtime=True)
    out = {"metric": "serving latency", "stream": {"requests": args.requests, "rate_per_s": args.rate,
           "seed": args.seed, "mix": "60% acoustic / 40% isotropic elastic, 50% forward / 50% gradient",
           "nt_range": [max(50, int(1000 * args.scale_nt)), max(60, int(6000 * args.scale_nt))],
           "shots_max": args.max_shots, "nz_max": args.max_nz, "nx_max": args.max_nx},
           "slots": args.slots, "serving": rep.summary()}
    print(f"[serving] done: p50 {out['serving']['latency_p50_s']:.3f} s, p99 {out['serving']['latency_p99_s']:.3f} s",
          file=sys.stderr, flush=True)
    if not args.no_plain:
        t0 = time.perf_counter()
        alone = run_alone(reqs)
        wall = time.perf_counter() - t0
        service = np.array([o.end - o.start for o in alone])
        out["plain_one_at_a_time"] = {"wall_s": wall, "throughput_req_per_s": len(alone) / wall,
                                      "service_p50_s": float(np.percentile(service, 50)),
                                      "service_p99_s": float(np.percentile(service, 99))}
        same = all(a.digest.keys() == b.digest.keys() and all(
            abs(a.digest[k][1] - b.digest[k][1]) <= 1e-6 * max(abs(b.digest[k][1]), 1e-30) for k in a.digest)
            for a, b in zip(rep.outcomes, alone) if a.error is None)
        out["digests_match_plain"] = bool(same)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
