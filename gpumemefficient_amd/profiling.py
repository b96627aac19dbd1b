"""Helpers behind the rocprofv3 summaries in profiles/.

    python -m gpumemefficient_amd.profiling copy [--bytes N]
        device-to-device copy of N bytes (default 1 GiB) through
        hipMemcpyDtoD: the calibration kernel for FETCH_SIZE / WRITE_SIZE
    python -m gpumemefficient_amd.profiling summarize --trace DIR [--pmc DIR ...] --out FILE
        kernel-trace statistics and per-dispatch counter values (rocprofv3
        --output-format csv) as a short Markdown summary
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


def summarize(trace: str | None, pmcs: list[str], out: str, title: str, command: str) -> None:
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", ""]
    if trace:
        lines += ["## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
                  "| kernel | calls | total ms | avg us | min us | max us |", "|---|---|---|---|---|---|"]
        for k, n, tot, avg, mn, mx in kernel_stats(trace):
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 30)
    s = sub.add_parser("summarize")
    s.add_argument("--trace")
    s.add_argument("--pmc", action="append", default=[])
    s.add_argument("--out", required=True)
    s.add_argument("--title", default="rocprofv3 summary")
    s.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_kernel(a.bytes)
    else:
        summarize(a.trace, a.pmc, a.out, a.title, a.command)
    return 0


if __name__ == "__main__":
    sys.exit(main())
