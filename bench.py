#!/usr/bin/env python3
"""Benchmark of the wavefront DP hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sw|ed|lcs|batch]

Default workload (the configuration the headline metric, % of int32 VALU
peak, is quoted on): Smith-Waterman (SW.cpp scoring +3/-3/-2) of two
131072-long 4-letter sequences, seed 3, with a planted 4096-long homologous
segment (10% substitutions) -- 1.72e10 DP cells -- computing what the
reference's SW() computes, the bottom-right cell; the same pair with this
library's running maximum as well is in extra.sw_131072_with_max.  A step is one pass of the
DP over the pair with its codes resident in HBM (one wf_plan_run: reset +
one persistent launch, synchronous).

One JSON line on stdout:
  value        GCUPS over the K timed steps (wall clock, whole job)
  roofline     the engine kernel against the int32 VALU peak of the chip
               (256 CUs x 64 lanes x max engine clock): achieved = algorithmic
               int ops per launch (SW 8, ED 7, LCS 4 per cell) / mean launch
               time from HIP events on the library's stream; traffic = HBM
               bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
               (profiles/traffic.json), when measured on this kernel variant
  cpu_baseline the oracle (C restatement of SW.cpp) timed on this host on the
               same pair: one thread over the whole pair, plus the all-cores
               anti-diagonal tile variant (thread count stated), plus the
               reference's CPU LCS program as it runs (the whole table) on
               its own 2^15 input through oracle/bin/LCS --faithful
  extra        the other measurements: edit distance 65536^2 and LCS 8192^2
               single pairs; the 4096-pair SW batch (one persistent launch,
               eager and replayed from a hipGraph, tile from the shipped
               autotune table) against the loop of one launch per pair; the
               reference's naive baselines (one launch per cell or tile
               anti-diagonal) on LCS 8192^2; the served Poisson stream (p50/p99
               latency and GCUPS per rate, continuous batching against one
               launch per request); the SAT and 2D-SOR table sweeps at
               2^15 x 2^15; the reference's published workload (LCS of its
               own 2^15 input file) through the drop-in LCS() call
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
    if name in ("sw", "sw_max"):
        # The reference's SW() returns the bottom-right cell only
        # (SmithWaterman/SW.cpp:80, GPU_Hyperlane_Share.cu:781): the workload
        # is that; "sw_max" also computes this library's running maximum.
        a, b = G.sw_pair()
        return [Job("SW", a, b, no_max=(name == "sw"))], "SW", {
            "workload": "smith_waterman_131072x131072_alpha4_seed3_planted4096", "n1": 131072, "n2": 131072,
            "scoring": "+3/-3/-2 (SmithWaterman/SW.cpp)",
            "outputs": ("bottom-right cell, as the reference's SW()" if name == "sw"
                        else "bottom-right cell and the running maximum")}
    if name == "lcs":
        a, b = G.lcs_pair()
        return [Job("LCS", a, b)], "LCS", {"workload": "lcs_8192x8192_alpha4_seed1", "n1": 8192, "n2": 8192}
    if name in ("batch", "batch_max"):
        pairs = G.batch_pairs()
        return [Job("SW", a, b, no_max=(name == "batch")) for a, b in pairs], "SW", {
            "workload": "smith_waterman_batch_4096_pairs_loguniform_64_4096_seed4", "pairs": 4096,
            "outputs": ("bottom-right cell per pair, as the reference's SW()" if name == "batch"
                        else "bottom-right cell and the running maximum per pair")}
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


# The reference's only published timing of this path: LCS of its own
# 32768 x 32768 input file (LCS/Data/str1_2_15_str2_2_15.txt, committed as
# tests/golden/data/) through the hyperlane Share host function, 13.6486 s
# for 100 calls on a GeForce GTX 1080 Ti, every call including the table
# allocation, the copies and the kernels (LCS/GPU_Hyperlane_Share.txt:103),
# printing lcs length 21469 (:102).
PUBLISHED_LCS_32768_MS_PER_CALL = 136.486
PUBLISHED_LCS_32768_LENGTH = 21469


def lcs_published_measure(calls: int = 20) -> dict:
    """The same input through the drop-in call the reference's driver makes,
    LCS(n1, n2, arr1, arr2, paddX, paddY, table) of libwfdp_compat_share.so
    (host arrays in, score out: encoding, upload, launch and result copy in
    every call), timed like the driver: wall clock of whole calls."""
    import ctypes as C
    from gpumemefficient_amd.datafile import read_input
    a, b = read_input(ROOT / "tests" / "golden" / "data" / "str1_2_15_str2_2_15.txt")
    lib = C.CDLL(str(ROOT / "gpumemefficient_amd" / "lib" / "libwfdp_compat_share.so"))
    fn = getattr(lib, "_Z3LCSiiPiS_iiS_")  # int LCS(int, int, int*, int*, int, int, int*)
    I = C.POINTER(C.c_int)
    fn.argtypes = [C.c_int, C.c_int, I, I, C.c_int, C.c_int, I]
    fn.restype = C.c_int
    pa, pb = a.ctypes.data_as(I), b.ctypes.data_as(I)
    length = fn(a.size, b.size, pa, pb, 3, 1, None)
    assert length == PUBLISHED_LCS_32768_LENGTH, length
    t0 = time.perf_counter()
    for _ in range(calls):
        length = fn(a.size, b.size, pa, pb, 3, 1, None)
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / calls
    cells = a.size * b.size
    return {"workload": "lcs_32768x32768_reference_input_LCS_Data_str1_2_15_str2_2_15", "calls": calls,
            "entry": "LCS(n1, n2, arr1, arr2, paddX, paddY, table) of libwfdp_compat_share.so",
            "ms_per_call_end_to_end": round(ms, 3), "gcups_end_to_end": round(cells / ms / 1e6, 1),
            "published_ms_per_call_gtx1080ti": PUBLISHED_LCS_32768_MS_PER_CALL,
            "speedup_vs_published": round(PUBLISHED_LCS_32768_MS_PER_CALL / ms, 1), "lcs_length": length}


def time_plan(plan, steps: int, warmup: int, per_pair: bool = False):
    for _ in range(warmup):
        plan.run(per_pair=per_pair)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kms.append(plan.run(per_pair=per_pair))   # synchronous: returns after the stream drains
    dt = time.perf_counter() - t0
    return dt, kms


def _host_cpu() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_threads() -> int:
    """Host threads this job may use: the box's share (OMP_NUM_THREADS when
    set), at most 16."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, min(16, n))


def cpu_baseline(jobs, prob: str, budget_s: float = 20.0, min_s: float = 10.0):
    """Time the oracle on the pair: one thread (like the reference programs)
    on its first rows (a band of the same table) if the whole pair would take
    longer than the budget, else on the whole pair, repeated until at least
    `min_s` of CPU work; then the all-cores variant (anti-diagonals of
    256 x 256 tiles on every thread this job may use) on the whole pair,
    repeated to at least `min_s` / 2."""
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
    threads = _cpu_threads()
    want = O.score(prob, a, b) if rows == b.size else None
    preps, pdt = 0, 0.0
    while preps == 0 or pdt < min_s / 2:
        t0 = time.perf_counter()
        got = O.blocked(prob, a, b, threads, 256)
        pdt += time.perf_counter() - t0
        preps += 1
    if want is not None:
        assert got == want, (got, want)
    pcells = a.size * b.size * preps
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port", "sample": sample,
            "host_cpu": _host_cpu(), "host_threads": os.cpu_count(),
            "all_cores": {"value": round(pcells / pdt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
                          "sample": (f"{prob} oracle over anti-diagonals of 256x256 tiles, {threads} threads, on the "
                                     f"whole {b.size}x{a.size} pair, {preps} run(s): {pcells:.3e} cells in "
                                     f"{pdt:.1f} s")},
            "faithful_lcs_32768": faithful_lcs_32768()}


def faithful_lcs_32768() -> dict:
    """The reference's CPU LCS program as it runs (the whole (n+1)^2 int32
    table, filled in its loop order, timed around the DP call with
    gettimeofday) on its own 2^15 input, through the drop-in CPU program
    oracle/bin/LCS --faithful; its score must be the published 21469."""
    import re
    import shutil
    import subprocess
    import tempfile
    exe = ROOT / "oracle" / "bin" / "LCS"
    src = ROOT / "tests" / "golden" / "data" / "str1_2_15_str2_2_15.txt"
    if not exe.exists() or not src.exists():
        return {"skipped": "oracle/bin/LCS not built" if not exe.exists() else "input file missing"}
    with tempfile.TemporaryDirectory() as td:
        (Path(td) / "Data").mkdir()
        shutil.copy(src, Path(td) / "Data" / src.name)
        out = subprocess.run([str(exe), "15", "15", "--faithful"], cwd=td, capture_output=True, text=True,
                             timeout=120).stdout
    length = int(re.search(r"lcs length: (\d+)", out).group(1))
    ms = float(re.search(r"execution time: ([0-9.e+-]+) ms", out).group(1))
    assert length == 21469, length
    return {"value": round(32768 * 32768 / ms / 1e6, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "ms": round(ms, 1), "lcs_length": length,
            "sample": ("oracle/bin/LCS 15 15 --faithful on the reference's LCS/Data/str1_2_15_str2_2_15.txt "
                       "(whole table, the reference program's memory behaviour), 1 run")}


def single_pair_extra(ctx, name: str, peak: float, steps: int = 3) -> dict:
    """One of the other single-pair workloads: kernel time, GCUPS and the
    fraction of the int32 VALU peak (algorithmic ops)."""
    jobs, prob, config = make_workload(name)
    p = ctx.plan(jobs)
    _, kms = time_plan(p, steps, 1)
    r = p.results()[0]
    k = sum(kms) / len(kms)
    cells = int(jobs[0].a.size) * int(jobs[0].b.size)
    out = {"workload": config["workload"], "gcups": round(cells / k / 1e6, 1), "kernel_ms": round(k, 3),
           "int_valu_peak_frac": round(OPS_PER_CELL[prob] * cells / (k * 1e-3) / 1e12 / peak, 4),
           "ops_per_cell": OPS_PER_CELL[prob], "last": r.last, "engine": p.describe()}
    if prob == "SW":
        out["max_score"] = r.max_score
    p.close()
    return out


def batch_extra(ctx) -> dict:
    """The 4096-pair SW batch: one persistent launch (eager, then replayed
    from a captured hipGraph) against the loop of one launch per pair, on the
    same plan; the planner's tile comes from the shipped autotune table."""
    bj, _, bc = make_workload("batch")
    p = ctx.plan(bj)
    _, bkms = time_plan(p, 5, 1)
    res_eager = [(r.last, r.max_score) for r in p.results()]
    p.capture()
    gdt, gkms = time_plan(p, 5, 1)
    res_graph = [(r.last, r.max_score) for r in p.results()]
    assert res_graph == res_eager
    _, pkms = time_plan(p, 3, 1, per_pair=True)
    assert [(r.last, r.max_score) for r in p.results()] == res_eager
    bcells = sum(int(j.a.size) * int(j.b.size) for j in bj)
    bk, gk, pk = sum(bkms) / len(bkms), sum(gkms) / len(gkms), sum(pkms) / len(pkms)
    # The same launch also computing every pair's running maximum.
    mj, _, _ = make_workload("batch_max")
    pm = ctx.plan(mj)
    _, mkms = time_plan(pm, 5, 1)
    assert [r.last for r in pm.results()] == [x[0] for x in res_eager]
    mk = sum(mkms) / len(mkms)
    pm.close()
    out = {"workload": bc["workload"], "outputs": bc["outputs"], "cells": bcells,
           "persistent_gcups": round(bcells / bk / 1e6, 1), "persistent_ms": round(bk, 3),
           "persistent_with_max_gcups": round(bcells / mk / 1e6, 1), "persistent_with_max_ms": round(mk, 3),
           "graph_replay_gcups": round(bcells / gk / 1e6, 1), "graph_replay_ms": round(gk, 3),
           "per_pair_launch_gcups": round(bcells / pk / 1e6, 1), "per_pair_launch_ms": round(pk, 3),
           "speedup": round(pk / min(bk, gk), 2), "engine": p.describe(), "tile_source": tile_source(ctx, bj)}
    p.close()
    return out


def tile_source(ctx, jobs) -> str:
    from gpumemefficient_amd import api
    t = api.tuned(jobs, ctx)
    return f"autotune table ({t.engine or 'default'} {t.strip_rows} rows, K={t.tile_w})" if t else "built-in defaults"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sw", choices=["sw", "ed", "lcs", "batch", "stencil"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the other single pairs, batch and serving")
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
        if args.workload == "sw":
            extra["sw_131072_with_max"] = single_pair_extra(ctx, "sw_max", peak)
        for name, key in (("sw", "sw_131072"), ("ed", "ed_65536"), ("lcs", "lcs_8192")):
            if args.workload != name:
                extra[key] = single_pair_extra(ctx, name, peak)
        if args.workload != "batch":
            extra["batch_4096"] = batch_extra(ctx)
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
        # lengths): continuous batching on stream slots, the resident
        # persistent launch, and one launch per request.
        from gpumemefficient_amd.serve_bench import run_rate
        serve = {"requests_per_rate": args.serve_requests, "continuous": [], "resident": [],
                 "per_request": []}
        for rate in (1000.0, 2000.0, 5000.0, 10000.0, 20000.0, 50000.0):
            serve["continuous"].append(run_rate(rate, args.serve_requests, False))
            serve["resident"].append(run_rate(rate, args.serve_requests, False, resident=True))
        serve["per_request"].append(run_rate(50000.0, args.serve_requests, True))
        sat_c = serve["continuous"][-1]["gcups"]
        sat_r = serve["resident"][-1]["gcups"]
        sat_p = serve["per_request"][-1]["gcups"]
        serve["saturated_speedup"] = round(sat_c / sat_p, 2) if sat_p else None
        serve["saturated_speedup_resident"] = round(sat_r / sat_p, 2) if sat_p else None
        extra["serving"] = serve
        extra["stencil_32768"] = stencil_measure(ctx)
        extra["lcs_32768_reference_input_drop_in"] = lcs_published_measure()
        line["extra"] = extra

    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(jobs, prob)
    ctx.close()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
