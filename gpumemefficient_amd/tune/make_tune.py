"""Write the autotune table shipped with the library (wfdp_tune.txt here,
which every context loads; WFDP_TUNE overrides or disables it).

    python -m gpumemefficient_amd.tune.make_tune [--out FILE] [--reps N]

Times every compiled kernel variant (wf_autotune) on the standard workloads
with their inputs resident in HBM and keeps the fastest per workload class:
the 4096-pair Smith-Waterman batch (lengths log-uniform in [64, 4096], seed
4), and the single pairs SW 131072^2, ED 65536^2 and LCS 8192^2 (SW as the
reference's SW() computes it: the bottom-right cell, no running maximum).  Needs a
GPU; the table is keyed by GPU architecture and CU count, so it only applies
to the same kind of device.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(HERE / "wfdp_tune.txt"))
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    from gpumemefficient_amd import Context, Job, api
    from gpumemefficient_amd import datagen as G
    api.tune_clear()
    # SW as the reference computes it: the bottom-right cell (no_max).
    workloads = {
        "sw_batch_4096": [Job("SW", a, b, no_max=True) for a, b in G.batch_pairs()],
        "sw_131072": [Job("SW", *G.sw_pair(), no_max=True)],
        "ed_65536": [Job("ED", *G.ed_pair())],
        "lcs_8192": [Job("LCS", *G.lcs_pair())],
    }
    with Context(0, timeout_s=30.0) as ctx:
        for name, jobs in workloads.items():
            tile, ms = ctx.autotune(jobs, reps=args.reps)
            print(json.dumps({"workload": name, "engine": tile.engine, "rows_per_lane": tile.rows_per_lane,
                              "strip_rows": tile.strip_rows, "chunk": tile.tile_w, "ms": round(ms, 4)}), flush=True)
    api.tune_save(args.out)
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
