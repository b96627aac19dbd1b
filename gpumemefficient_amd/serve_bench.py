"""Serving benchmark: a Poisson request stream of 50% edit distance
(20-letter) and 50% Smith-Waterman (4-letter) requests, lengths 80% in
[100, 1000], 18% in [1000, 16384], 2% in [16384, 131072] (seed 5), replayed
in real time against the continuous-batching server (launches on stream
slots, or one resident persistent launch) and against the
one-launch-per-request baseline.  Reports p50/p99 latency (completion seen
by the host minus scheduled arrival) and GCUPS (cells / makespan) per rate.

    python -m gpumemefficient_amd.serve_bench [--rates 1000,2000,...] [--requests N] [--resident] [--baseline]
"""
from __future__ import annotations

import argparse
import json
import sys

from .api import Job
from .datagen import serving_requests
from .server import Server, percentile


def run_rate(rate: float, count: int, per_request: bool, seed: int = 5, resident: bool = False) -> dict:
    reqs = serving_requests(rate, count, seed=seed)
    jobs = [Job(r.problem, r.a, r.b) for r in reqs]
    arrivals = [r.arrival_s for r in reqs]
    cells = sum(r.cells for r in reqs)
    with Server(0, per_request=per_request, resident=resident) as srv:
        comps = srv.replay(jobs, arrivals)
        st = srv.stats()
    bad = sum(1 for c in comps if c.status != 0)
    lat = [c.latency for c in comps]
    makespan = max(c.t_done for c in comps) - min(arrivals)
    return {"rate_req_s": rate, "requests": count, "cells": cells, "errors": bad,
            "p50_ms": round(percentile(lat, 50) * 1e3, 3), "p99_ms": round(percentile(lat, 99) * 1e3, 3),
            "gcups": round(cells / makespan / 1e9, 1), "makespan_s": round(makespan, 4),
            "launches": int(st["launches"]), "offered_gcups": round(cells / max(arrivals) / 1e9, 1)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="1000,2000,5000,10000,20000,50000")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--baseline", action="store_true", help="also run one launch per request")
    ap.add_argument("--resident", action="store_true", help="also run the resident (persistent-launch) server")
    args = ap.parse_args(argv)
    out = []
    for r in [float(x) for x in args.rates.split(",")]:
        row = {"mode": "continuous", **run_rate(r, args.requests, False)}
        print(json.dumps(row), flush=True)
        out.append(row)
        if args.resident:
            row = {"mode": "resident", **run_rate(r, args.requests, False, resident=True)}
            print(json.dumps(row), flush=True)
            out.append(row)
        if args.baseline:
            row = {"mode": "per_request", **run_rate(r, args.requests, True)}
            print(json.dumps(row), flush=True)
            out.append(row)
    return 0


if __name__ == "__main__":
    sys.exit(main())
