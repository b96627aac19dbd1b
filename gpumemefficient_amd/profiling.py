"""Helpers behind the rocprofv3 summaries in profiles/.

    python -m gpumemefficient_amd.profiling copy [--bytes N]
        device-to-device copy of N bytes (default 1 GiB) through
        hipMemcpyDtoD: the calibration kernel for FETCH_SIZE / WRITE_SIZE
    python -m gpumemefficient_amd.profiling summarize --trace DIR [--pmc DIR ...] --out FILE
        kernel-trace statistics and per-dispatch counter values (rocprofv3
        --output-format csv) as a short Markdown summary
    python -m gpumemefficient_amd.profiling metrics --trace DIR --pmc DIR ... --kernel SUBSTR
            --problem SW --n1 N --n2 N --strip-rows R --ref-gld P --ref-gst P ... --out FILE
        the reference's nvprof metrics (SmithWaterman/script.sh:3) next to
        their rocprofv3 counterparts for one strip-engine kernel
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import glob
import os
import sys
from collections import defaultdict


def copy_kernel(nbytes: int) -> None:
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), nbytes) or hip.hipMalloc(C.byref(dst), nbytes):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, nbytes)
    hip.hipDeviceSynchronize()
    hipMemcpyDeviceToDevice = 3
    for _ in range(3):
        if hip.hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToDevice):
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copied {nbytes} bytes device-to-device 3 times")


def _read_csv(pattern: str):
    rows = []
    for f in glob.glob(pattern, recursive=True):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def kernel_stats(trace_dir: str):
    """(name, calls, total_ns, avg_ns, min_ns, max_ns) per kernel."""
    rows = _read_csv(os.path.join(trace_dir, "**", "*kernel_trace.csv"))
    per = defaultdict(list)
    for r in rows:
        per[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = []
    for k, v in per.items():
        out.append((k, len(v), sum(v), sum(v) / len(v), min(v), max(v)))
    out.sort(key=lambda t: -t[2])
    return out


def counters(pmc_dir: str):
    """{kernel: {counter: mean value per dispatch}}"""
    rows = _read_csv(os.path.join(pmc_dir, "**", "*counter_collection.csv"))
    acc = defaultdict(lambda: defaultdict(list))
    for r in rows:
        acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


def summarize(traces: list[str], pmcs: list[str], out: str, title: str, command: str, note: str = "") -> None:
    """Markdown summary of kernel traces and counter runs.  A trace given as
    LABEL=DIR gets a section titled with its label."""
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", ""]
    if note:
        lines += [note, ""]
    for t in traces:
        label, _, d = t.rpartition("=")
        head = f"## Kernel trace: {label}" if label else "## Kernel trace (rocprofv3 --kernel-trace --stats)"
        lines += [head, "", "| kernel | calls | total ms | avg us | min us | max us |", "|---|---|---|---|---|---|"]
        for k, n, tot, avg, mn, mx in kernel_stats(d):
            lines.append(f"| `{k[:90]}` | {n} | {tot / 1e6:.3f} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} |")
        lines.append("")
    for d in pmcs:
        lines += [f"## Counters ({os.path.basename(d.rstrip('/'))})", "", "| kernel | counter | mean per dispatch |",
                  "|---|---|---|"]
        for k, cs in counters(d).items():
            for c, v in sorted(cs.items()):
                lines.append(f"| `{k[:70]}` | {c} | {v:.6g} |")
        lines.append("")
    with open(out, "w") as fh:
        fh.write("\n".join(lines))
    print(f"wrote {out}")


# HBM counter scales from the 1 GiB device-to-device copy: a copy dispatch
# reads 524300 KiB of FETCH_SIZE and 1048576 KiB of WRITE_SIZE per GiB moved.
FETCH_SCALE = 1.9999
WRITE_SCALE = 1.0
OPS_PER_CELL = {"LCS": 4, "ED": 7, "SW": 8}


def strip_bytes(n1: int, n2: int, strip_rows: int) -> tuple[int, int]:
    """Algorithmic global-memory bytes of one strip-engine launch: every
    strip boundary row (n1 eight-byte tagged words) is stored once and
    loaded once; every strip loads the n1 one-byte codes of a, and the n2
    codes of b are loaded once."""
    strips = (n2 + strip_rows - 1) // strip_rows
    boundary = (strips - 1) * n1 * 8
    return boundary + strips * n1 + n2, boundary


def split_widths(n1: int) -> list[int]:
    """Column counts of the three pairs an SW column split launches: the
    left columns 1..n1/2, the right columns from a zero boundary and the
    reversed right columns."""
    mid = n1 // 2
    return [mid, n1 - mid, n1 - mid]


def split_bytes(n1: int, n2: int, strip_rows: int) -> tuple[int, int]:
    """strip_bytes of the three pairs of an SW column split, plus the last
    column (n2 four-byte words) that the left and the reversed pair store
    for the combine kernel."""
    load_b = store_b = 0
    for w in split_widths(n1):
        lb, sb = strip_bytes(w, n2, strip_rows)
        load_b, store_b = load_b + lb, store_b + sb
    return load_b, store_b + 2 * n2 * 4


def metrics(trace: str, pmcs: list[str], kernel: str, problem: str, n1: int, n2: int, strip_rows: int,
            cus: int, clock_mhz: float, ref: dict, title: str, command: str, out: str,
            column_split: bool = False) -> None:
    """rocprofv3 counterparts of the reference's nvprof metrics for the
    dispatches of `kernel` (mean per dispatch):
      gld_efficiency       algorithmic load bytes / HBM bytes fetched
      gst_efficiency       algorithmic store bytes / HBM bytes written
      shared_efficiency    1 - SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE
      tex_cache_hit_rate   1 - TCP_TCC_READ_REQ / TCP_TOTAL_CACHE_ACCESSES (vL1D)
      l2_tex_read_hit_rate TCC_HIT / (TCC_HIT + TCC_MISS)
      inst_integer         SQ_INSTS_VALU_INT32 (wave instructions) x 64 lanes
    plus the VALU issue utilisation and the int32 VALU rate against the
    chip's peak.  With column_split the dispatch is the three pairs of an
    SW column split (split_widths): 1.5x the cells of the pair."""
    c = {}
    for d in pmcs:
        for k, cs in counters(d).items():
            if kernel in k:
                c.update(cs)
    stats = [t for t in kernel_stats(trace) if kernel in t[0]]
    if not stats:
        raise SystemExit(f"no dispatch of {kernel} in {trace}")
    name, calls, _, avg_ns, _, _ = stats[0]
    load_b, store_b = (split_bytes if column_split else strip_bytes)(n1, n2, strip_rows)
    cells = n2 * (sum(split_widths(n1)) if column_split else n1)
    fetch = c.get("FETCH_SIZE", 0.0) * 1024 * FETCH_SCALE
    write = c.get("WRITE_SIZE", 0.0) * 1024 * WRITE_SCALE
    peak = cus * 64 * clock_mhz * 1e6
    int_ops = c.get("SQ_INSTS_VALU_INT32", 0.0) * 64
    rows = []

    def pct(x):
        return f"{100 * x:.2f}%"

    rows.append(("gld_efficiency", ref.get("gld"), "algorithmic load bytes / (FETCH_SIZE x 2.0)",
                 f"{load_b:.4g} B / {fetch:.4g} B = {pct(load_b / fetch)}" if fetch else "n/a"))
    rows.append(("gst_efficiency", ref.get("gst"), "algorithmic store bytes / (WRITE_SIZE x 1.0)",
                 f"{store_b:.4g} B / {write:.4g} B = {pct(store_b / write)}" if write else "n/a"))
    if "SQ_LDS_IDX_ACTIVE" in c and c["SQ_LDS_IDX_ACTIVE"] > 0:
        rows.append(("shared_efficiency", ref.get("shared"), "1 - SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE",
                     pct(1 - c.get("SQ_LDS_BANK_CONFLICT", 0.0) / c["SQ_LDS_IDX_ACTIVE"])))
    if c.get("TCP_TOTAL_CACHE_ACCESSES_sum"):
        rows.append(("tex_cache_hit_rate", ref.get("tex"), "1 - TCP_TCC_READ_REQ / TCP_TOTAL_CACHE_ACCESSES (vL1D)",
                     pct(1 - c.get("TCP_TCC_READ_REQ_sum", 0.0) / c["TCP_TOTAL_CACHE_ACCESSES_sum"])))
    hm = c.get("TCC_HIT_sum", 0.0) + c.get("TCC_MISS_sum", 0.0)
    if hm:
        rows.append(("l2_tex_read_hit_rate", ref.get("l2"), "TCC_HIT / (TCC_HIT + TCC_MISS)",
                     pct(c["TCC_HIT_sum"] / hm)))
    if int_ops:
        rows.append(("inst_integer", ref.get("inst_integer"), "SQ_INSTS_VALU_INT32 x 64 (lane ops)",
                     f"{int_ops:.4g}"))
    lines = [f"# {title}", "", "Commands:", "", "```", command, "```", "",
             f"Kernel `{name}`: {calls} dispatch(es), mean {avg_ns / 1e3:.1f} us.", "",
             "| nvprof metric | reference (GTX 1080 Ti) | rocprofv3 counterpart | MI355X |", "|---|---|---|---|"]
    for m, r, how, v in rows:
        lines.append(f"| {m} | {r if r is not None else '-'} | {how} | {v} |")
    lines.append("")
    extra = []
    if c.get("SQ_INSTS_VALU"):
        # A wave64 VALU instruction occupies its SIMD for 4 cycles.
        simd_cycles = 4 * cus * avg_ns * 1e-9 * clock_mhz * 1e6
        extra.append(f"- VALU issue utilisation (SQ_INSTS_VALU x 4 cycles / (4 SIMDs x {cus} CUs x kernel "
                     f"cycles at {clock_mhz:.0f} MHz)): {100 * c['SQ_INSTS_VALU'] * 4 / simd_cycles:.1f}%")
    if int_ops:
        extra.append(f"- int32 VALU lane ops per dispatch {int_ops:.4g} over {avg_ns / 1e3:.1f} us = "
                     f"{int_ops / (avg_ns * 1e-9) / 1e12:.2f} TOP/s, {100 * int_ops / (avg_ns * 1e-9) / peak:.1f}% "
                     f"of the {peak / 1e12:.2f} TOP/s int32 VALU peak ({cus} CUs x 64 lanes x {clock_mhz:.0f} MHz)")
        extra.append(f"- counted int32 lane ops per cell: {int_ops / cells:.2f} (algorithmic: "
                     f"{OPS_PER_CELL[problem]}); all VALU instructions per cell: "
                     f"{c.get('SQ_INSTS_VALU', 0.0) * 64 / cells:.2f}")
    if fetch or write:
        extra.append(f"- HBM bytes per dispatch: fetched {fetch:.4g}, written {write:.4g}; algorithmic "
                     f"{load_b + store_b:.4g} ({n2} rows in strips of {strip_rows}, n1 = {n1}"
                     + (f", column split into pairs of {split_widths(n1)} columns, {cells} cells)"
                        if column_split else ")"))
    lines += extra + ["", "Raw counters (mean per dispatch):", "", "| counter | value |", "|---|---|"]
    for k in sorted(c):
        lines.append(f"| {k} | {c[k]:.6g} |")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(f"wrote {out}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 30)
    s = sub.add_parser("summarize")
    s.add_argument("--trace", action="append", default=[], help="DIR or LABEL=DIR; may repeat")
    s.add_argument("--pmc", action="append", default=[])
    s.add_argument("--out", required=True)
    s.add_argument("--title", default="rocprofv3 summary")
    s.add_argument("--command", default="")
    s.add_argument("--note", default="", help="a paragraph placed after the command")
    m = sub.add_parser("metrics")
    m.add_argument("--trace", required=True)
    m.add_argument("--pmc", action="append", default=[])
    m.add_argument("--kernel", required=True, help="substring of the kernel name")
    m.add_argument("--problem", required=True, choices=sorted(OPS_PER_CELL))
    m.add_argument("--n1", type=int, required=True)
    m.add_argument("--n2", type=int, required=True)
    m.add_argument("--strip-rows", type=int, required=True)
    m.add_argument("--column-split", action="store_true",
                   help="the dispatch runs an SW pair as three column pairs")
    m.add_argument("--cus", type=int, default=256)
    m.add_argument("--clock-mhz", type=float, default=2400.0)
    for key in ("gld", "gst", "shared", "tex", "l2", "inst-integer"):
        m.add_argument(f"--ref-{key}", help="the reference's nvprof value, quoted as printed")
    m.add_argument("--title", default="nvprof metrics and their rocprofv3 counterparts")
    m.add_argument("--command", default="")
    m.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_kernel(a.bytes)
    elif a.cmd == "metrics":
        ref = {"gld": a.ref_gld, "gst": a.ref_gst, "shared": a.ref_shared, "tex": a.ref_tex, "l2": a.ref_l2,
               "inst_integer": a.ref_inst_integer}
        metrics(a.trace, a.pmc, a.kernel, a.problem, a.n1, a.n2, a.strip_rows, a.cus, a.clock_mhz, ref, a.title,
                a.command, a.out, a.column_split)
    else:
        summarize(a.trace, a.pmc, a.out, a.title, a.command, a.note)
    return 0


if __name__ == "__main__":
    sys.exit(main())
