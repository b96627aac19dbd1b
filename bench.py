#!/usr/bin/env python3
"""Benchmark of the wavefront DP hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ed|sw|lcs|batch]

Default workload (the configuration the headline metric is quoted on that
fits one GPU): edit distance (ED.cpp semantics) between two i.i.d. uniform
20-letter sequences of 65536 symbols, seed 2 -- 4.29e9 DP cells.  A step is
one pass of the DP over the pair with its codes resident in HBM (one
wf_plan_run: reset + one persistent launch, synchronous).

One JSON line on stdout:
  value        GCUPS over the K timed steps (wall clock, whole job)
  roofline     the engine kernel against the int32 VALU peak of the chip
               (256 CUs x 64 lanes x max engine clock): achieved = algorithmic
               int ops per launch (ED 7, SW 8, LCS 4 per cell) / mean launch
               time from HIP events on the library's stream
  cpu_baseline the oracle (C restatement of ED.cpp, one thread) timed on this
               host on the same pair
  extra        the other headline measurements: Smith-Waterman 131072^2
               single pair; the 4096-pair batch against the loop of one
               launch per pair; the reference's naive baselines (one launch
               per cell or tile anti-diagonal) on LCS 8192^2; the served
               Poisson stream (p50/p99 latency and GCUPS per rate, continuous
               batching against one launch per request); the SAT and 2D-SOR
               table sweeps at 2^15 x 2^15 (the reference's nvprof runs of
               those programs) against the HBM bandwidth they are bound by
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

OPS_PER_CELL = {"LCS": 4, "ED": 7, "SW": 8}
PROFILES = ROOT / "profiles"


def peak_int32_tops(info: dict) -> float:
    """int32 VALU ops/s of the chip: one 64-lane op per CU per cycle."""
    return info["compute_units"] * 64 * info["max_clock_mhz"] * 1e6 / 1e12


def kernel_name(desc: dict) -> str:
    if desc["engine"] == "bitpar":
        return f"wf_engine_bitpar<{desc['waves']},{desc['chunk']}>"
    if desc["engine"] == "bitpar_staged":
        return f"wf_engine_bitpar_staged<{desc['waves']}>"
    base = "wf_engine" if desc["engine"] == "lockstep" else "wf_engine_async"
    return f"{base}<{desc['rows_per_lane']},{desc['waves']},{desc['chunk']}>"


def traffic_for(workload: str, desc: dict):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), as recorded in
    profiles/traffic.json -- only if it was measured on the kernel variant
    this run uses."""
    f = PROFILES / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text()).get(workload, {})
    except (OSError, ValueError):
        return None
    if d.get("kernel") != kernel_name(desc):
        return None
    return d.get("hbm_bytes_per_launch")


def make_workload(name: str):
    from gpumemefficient_amd import Job
    from gpumemefficient_amd import datagen as G
    if name == "ed":
        a, b = G.ed_pair()
        return [Job("ED", a, b)], "ED", {"workload": "edit_distance_65536x65536_alpha20_seed2", "n1": 65536,
                                         "n2": 65536, "alphabet": 20, "seed": 2, "semantics": "EditDistance/ED.cpp"}
    if name == "sw":
        a, b = G.sw_pair()
        return [Job("SW", a, b)], "SW", {"workload": "smith_waterman_131072x131072_alpha4_seed3_planted4096",
                                         "n1": 131072, "n2": 131072, "scoring": "+3/-3/-2 (SmithWaterman/SW.cpp)"}
    if name == "lcs":
        a, b = G.lcs_pair()
        return [Job("LCS", a, b)], "LCS", {"workload": "lcs_8192x8192_alpha4_seed1", "n1": 8192, "n2": 8192}
    if name == "batch":
        pairs = G.batch_pairs()
        return [Job("SW", a, b) for a, b in pairs], "SW", {
            "workload": "smith_waterman_batch_4096_pairs_loguniform_64_4096_seed4", "pairs": 4096}
    raise SystemExit(f"unknown workload {name}")


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def stencil_measure(ctx, exp: int = 15, steps: int = 3, warmup: int = 1, seed: int = 6) -> dict:
    """SAT and 2D-SOR sweeps of a (2^exp + 2) x (2^exp + 3) table of uniform
    [0, 10000) values (the reference generator's range), resident in HBM.
    Algorithmic HBM traffic: every interior cell read once and written once,
    8 bytes per cell."""
    import numpy as np
    from gpumemefficient_amd._hip import DeviceBuffer
    n = 1 << exp
    ld = (n + 3 + 3) & ~3
    host = np.random.default_rng(seed).integers(0, 10000, size=(n + 2, ld), dtype=np.int32)
    dev = DeviceBuffer.from_array(host)
    out = {"workload": f"table_{n}x{n}_uniform10000_seed{seed}", "n1": n, "n2": n, "bytes_per_cell": 8}
    for op in ("SAT", "SOR"):
        dev.upload(host)
        for _ in range(warmup):
            ctx.stencil_device(op, dev.ptr, n, n, ld)
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            kms.append(ctx.stencil_device(op, dev.ptr, n, n, ld).kernel_ms)
        dt = time.perf_counter() - t0
        k = sum(kms) / len(kms)
        cells = n * n
        gbs = 8 * cells / (k * 1e-3) / 1e9
        out[op.lower()] = {"kernel_ms": round(k, 3), "gcups": round(cells / k / 1e6, 1),
                           "wall_gcups": round(cells * steps / dt / 1e9, 1),
                           "hbm_gbs": round(gbs, 1), "hbm_peak_frac": round(gbs / HBM_PEAK_GBS, 4)}
    dev.free()
    return out


# The reference's only published timing of this path: LCS 32768 x 32768
# (4 letters) through the hyperlane Share host function, 13.6486 s for 100
# calls on a GeForce GTX 1080 Ti, every call including the table allocation,
# the copies and the kernels (LCS/GPU_Hyperlane_Share.txt:103).
PUBLISHED_LCS_32768_MS_PER_CALL = 136.486


def lcs_published_measure(ctx, calls: int = 20, seed: int = 7) -> dict:
    """The same workload through the drop-in path (host arrays in, score
    out: symbol encoding, upload, launch and result copy per call), timed
    like the reference's driver: wall clock of whole calls."""
    from gpumemefficient_amd import datagen as G
    a, b = G.uniform_pair(32768, 32768, 4, seed)
    ctx.align("LCS", a, b)
    t0 = time.perf_counter()
    kms = []
    for _ in range(calls):
        r = ctx.align("LCS", a, b)
        kms.append(r.kernel_ms)
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / calls
    cells = 32768 * 32768
    return {"workload": f"lcs_32768x32768_alpha4_seed{seed}", "calls": calls,
            "ms_per_call_end_to_end": round(ms, 3), "gcups_end_to_end": round(cells / ms / 1e6, 1),
            "kernel_ms": round(sum(kms) / len(kms), 3),
            "published_ms_per_call_gtx1080ti": PUBLISHED_LCS_32768_MS_PER_CALL,
            "speedup_vs_published": round(PUBLISHED_LCS_32768_MS_PER_CALL / ms, 1), "lcs_length": r.last}


def time_plan(plan, steps: int, warmup: int, per_pair: bool = False):
    for _ in range(warmup):
        plan.run(per_pair=per_pair)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kms.append(plan.run(per_pair=per_pair))   # synchronous: returns after the stream drains
    dt = time.perf_counter() - t0
    return dt, kms


def cpu_baseline(jobs, prob: str, budget_s: float = 20.0, min_s: float = 10.0):
    """Time the oracle (single thread) on the pair: on its first rows (a band
    of the same table) if the whole pair would take longer than the budget,
    else on the whole pair, repeated until at least `min_s` of CPU work."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    a, b = jobs[0].a, jobs[0].b
    # Probe the rate on a small band, then size the sample to the budget.
    probe_rows = max(1, min(b.size, 256))
    t0 = time.perf_counter()
    O.score(prob, a, b[:probe_rows])
    rate = a.size * probe_rows / max(1e-9, time.perf_counter() - t0)
    rows = int(min(b.size, max(probe_rows, budget_s * rate / a.size)))
    reps, dt = 0, 0.0
    while reps == 0 or (rows == b.size and dt < min_s):
        t0 = time.perf_counter()
        O.score(prob, a, b[:rows])
        dt += time.perf_counter() - t0
        reps += 1
    cells = a.size * rows * reps
    what = "the whole" if rows == b.size else f"rows 0..{rows - 1} of the"
    sample = (f"{prob} oracle (C restatement, -O3, 1 thread) on {what} {b.size}x{a.size} pair, "
              f"{reps} run(s): {cells:.3e} cells in {dt:.1f} s")
    host = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                host = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port", "sample": sample,
            "host_cpu": host, "host_threads": os.cpu_count()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ed", choices=["ed", "sw", "lcs", "batch", "stencil"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the SW single-pair, batch and serving measurements")
    ap.add_argument("--serve-requests", type=int, default=600, help="requests per rate in the serving measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(one huge pair would shard as bands of rows, each GPU streaming its bottom boundary row to the next)",
              file=sys.stderr)
        return 2

    from gpumemefficient_amd import Context
    ctx = Context(0, timeout_s=30.0)
    info = ctx.info()
    peak = peak_int32_tops(info)
    if args.workload == "stencil":
        # The SAT and 2D-SOR sweeps alone (for profiling); one JSON line.
        print(json.dumps({"stencil_32768": stencil_measure(ctx, steps=args.steps, warmup=args.warmup)}))
        ctx.close()
        return 0

    jobs, prob, config = make_workload(args.workload)
    cells = sum(int(j.a.size) * int(j.b.size) for j in jobs)
    plan = ctx.plan(jobs)
    dt, kms = time_plan(plan, args.steps, args.warmup)
    res = plan.results()
    desc = plan.describe()
    kernel_ms = sum(kms) / len(kms)
    gcups = cells * args.steps / dt / 1e9
    achieved = OPS_PER_CELL[prob] * cells / (kernel_ms * 1e-3) / 1e12
    config.update({"engine": desc["engine"], "rows_per_lane": desc["rows_per_lane"], "waves": desc["waves"],
                   "chunk": desc["chunk"], "strips": desc["tasks"], "grid": desc["grid"]})
    if desc["engine"].startswith("bitpar"):
        # A split pair runs as a forward and a reversed column half: twice
        # the strips of the unsplit pair.
        strip_rows = 64 * desc["waves"] * desc["rows_per_lane"]
        unsplit = sum((int(j.b.size) + strip_rows - 1) // strip_rows for j in jobs if j.a.size and j.b.size)
        config["column_split"] = desc["tasks"] > unsplit
    line = {
        "metric": "GCUPS & % int-VALU peak; batched vs one-pair-per-launch; p50/p99 latency",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded PCG64, gpumemefficient_amd.datagen)",
        "config": config,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(peak, 3),
            "unit": "TOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic_for(args.workload, desc),
            "kernel": kernel_name(desc),
            "ops_per_cell": OPS_PER_CELL[prob],
            "kernel_ms": round(kernel_ms, 4),
        },
        "result": {"last": res[0].last, "max_score": res[0].max_score} if len(res) == 1 else
                  {"pairs": len(res), "sum_last": int(sum(r.last for r in res))},
    }
    plan.close()

    if not args.no_extra:
        extra = {}
        if args.workload != "sw":
            sj, sp, sc = make_workload("sw")
            p = ctx.plan(sj)
            sdt, skms = time_plan(p, 3, 1)
            sr = p.results()[0]
            sk = sum(skms) / len(skms)
            scells = int(sj[0].a.size) * int(sj[0].b.size)
            extra["sw_131072"] = {
                "gcups": round(scells / sk / 1e6, 1), "kernel_ms": round(sk, 3),
                "int_valu_peak_frac_8ops": round(8 * scells / (sk * 1e-3) / 1e12 / peak, 4),
                "last": sr.last, "max_score": sr.max_score, "engine": p.describe()}
            p.close()
        if args.workload != "batch":
            bj, bp, bc = make_workload("batch")
            p = ctx.plan(bj)
            bdt, bkms = time_plan(p, 5, 1)
            pdt, pkms = time_plan(p, 3, 1, per_pair=True)
            bcells = sum(int(j.a.size) * int(j.b.size) for j in bj)
            bk, pk = sum(bkms) / len(bkms), sum(pkms) / len(pkms)
            extra["batch_4096"] = {
                "cells": bcells, "persistent_gcups": round(bcells / bk / 1e6, 1), "persistent_ms": round(bk, 3),
                "per_pair_launch_gcups": round(bcells / pk / 1e6, 1), "per_pair_launch_ms": round(pk, 3),
                "speedup": round(pk / bk, 2), "engine": p.describe()}
            p.close()
        # The reference's naive GPU baselines on the LCS 8192 x 8192 pair:
        # one launch per cell anti-diagonal, one launch per tile
        # anti-diagonal, against the strip engine on the same pair.
        lj, _, lc = make_workload("lcs")
        p = ctx.plan(lj)
        _, lkms = time_plan(p, 5, 1)
        lk = sum(lkms) / len(lkms)
        lr = p.results()[0]
        lcells = int(lj[0].a.size) * int(lj[0].b.size)
        naive = {"workload": lc["workload"], "engine_ms": round(lk, 3), "engine_gcups": round(lcells / lk / 1e6, 1)}
        for method in ("diagonal", "tile"):
            ctx.align_naive("LCS", lj[0].a, lj[0].b, method=method)
            nr, _ = ctx.align_naive("LCS", lj[0].a, lj[0].b, method=method)
            assert nr.last == lr.last, (method, nr.last, lr.last)
            naive[f"naive_{method}_ms"] = round(nr.kernel_ms, 3)
            naive[f"naive_{method}_gcups"] = round(lcells / nr.kernel_ms / 1e6, 2)
            naive[f"speedup_vs_{method}"] = round(nr.kernel_ms / lk, 1)
        naive["lcs_length"] = lr.last
        extra["vs_naive_lcs_8192"] = naive
        p.close()
        # Served mixed workload (Poisson arrivals, 50% ED / 50% SW, mixed
        # lengths), continuous batching against one launch per request.
        from gpumemefficient_amd.serve_bench import run_rate
        serve = {"requests_per_rate": args.serve_requests, "continuous": [], "per_request": []}
        for rate in (1000.0, 2000.0, 5000.0, 10000.0, 20000.0, 50000.0):
            serve["continuous"].append(run_rate(rate, args.serve_requests, False))
        serve["per_request"].append(run_rate(50000.0, args.serve_requests, True))
        sat_c = serve["continuous"][-1]["gcups"]
        sat_p = serve["per_request"][-1]["gcups"]
        serve["saturated_speedup"] = round(sat_c / sat_p, 2) if sat_p else None
        extra["serving"] = serve
        extra["stencil_32768"] = stencil_measure(ctx)
        extra["lcs_32768_drop_in"] = lcs_published_measure(ctx)
        line["extra"] = extra

    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(jobs, prob)
    ctx.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
