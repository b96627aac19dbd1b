"""The rocprofv3 summary tool (gpumemefficient_amd/profiling.py) on small
hand-written trace and counter files: labelled kernel-trace sections, the
counter tables, and the nvprof-metric counterparts computed from them."""
import csv

from gpumemefficient_amd import profiling as P


def _write(path, header, rows):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def _trace(d, spans):
    _write(d / "t_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [(k, s, e) for k, s, e in spans])


def _pmc(d, vals):
    _write(d / "a_counter_collection.csv", ["Kernel_Name", "Counter_Name", "Counter_Value"],
           [(k, c, v) for k, c, v in vals])


def test_kernel_stats_and_counters(tmp_path):
    _trace(tmp_path / "tr", [("k1", 0, 1000), ("k1", 5000, 8000), ("k2", 0, 100)])
    st = {t[0]: t[1:] for t in P.kernel_stats(str(tmp_path / "tr"))}
    assert st["k1"] == (2, 4000, 2000.0, 1000, 3000)
    assert st["k2"] == (1, 100, 100.0, 100, 100)
    _pmc(tmp_path / "pm", [("k1", "FETCH_SIZE", 10), ("k1", "FETCH_SIZE", 30), ("k2", "WRITE_SIZE", 7)])
    c = P.counters(str(tmp_path / "pm"))
    assert c["k1"]["FETCH_SIZE"] == 20 and c["k2"]["WRITE_SIZE"] == 7


def test_summarize_labelled_traces(tmp_path):
    _trace(tmp_path / "a", [("kern_a", 0, 2000)])
    _trace(tmp_path / "b", [("kern_b", 0, 3000)])
    _pmc(tmp_path / "p", [("kern_a", "SQ_WAVES", 64)])
    out = tmp_path / "s.md"
    P.main(["summarize", "--trace", f"first run={tmp_path / 'a'}", "--trace", str(tmp_path / "b"),
            "--pmc", str(tmp_path / "p"), "--title", "T", "--command", "cmd line", "--note", "a note",
            "--out", str(out)])
    text = out.read_text()
    assert text.startswith("# T\n") and "cmd line" in text and "a note" in text
    assert "## Kernel trace: first run" in text and "## Kernel trace (rocprofv3" in text
    assert "| `kern_a` | 1 | 0.002 | 2.0 |" in text and "| `kern_b` | 1 | 0.003 | 3.0 |" in text
    assert "| `kern_a` | SQ_WAVES | 64 |" in text


def test_metrics_counterparts(tmp_path):
    # One dispatch of 1 ms; 2 strips of 512 rows over n1 = 1024 columns.
    _trace(tmp_path / "t", [("wf_engine_async<2, 4, 32>", 0, 1_000_000)])
    fetch_kib, write_kib = 20.0, 10.0
    _pmc(tmp_path / "c", [("wf_engine_async<2, 4, 32>", "FETCH_SIZE", fetch_kib),
                          ("wf_engine_async<2, 4, 32>", "WRITE_SIZE", write_kib),
                          ("wf_engine_async<2, 4, 32>", "SQ_LDS_BANK_CONFLICT", 25),
                          ("wf_engine_async<2, 4, 32>", "SQ_LDS_IDX_ACTIVE", 100)])
    out = tmp_path / "m.md"
    P.main(["metrics", "--trace", str(tmp_path / "t"), "--pmc", str(tmp_path / "c"), "--kernel", "wf_engine_async",
            "--problem", "SW", "--n1", "1024", "--n2", "1024", "--strip-rows", "512", "--ref-gld", "86.30%",
            "--out", str(out)])
    text = out.read_text()
    load_b, store_b = P.strip_bytes(1024, 1024, 512)
    # One boundary row of 1024 eight-byte words, stored once and loaded once.
    assert store_b == 1024 * 8 and load_b == 1024 * 8 + 2 * 1024 + 1024
    gld = load_b / (fetch_kib * 1024 * P.FETCH_SCALE)
    gst = store_b / (write_kib * 1024 * P.WRITE_SCALE)
    assert f"{100 * gld:.2f}%" in text and f"{100 * gst:.2f}%" in text
    assert "| shared_efficiency | - | 1 - SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE | 75.00% |" in text
    assert "| gld_efficiency | 86.30% |" in text


def test_metrics_column_split(tmp_path):
    # An SW pair of 1025 columns splits into pairs of 512, 513 and 513.
    _trace(tmp_path / "t", [("wf_engine_async<2, 4, 32>", 0, 1_000_000)])
    _pmc(tmp_path / "c", [("wf_engine_async<2, 4, 32>", "FETCH_SIZE", 20.0),
                          ("wf_engine_async<2, 4, 32>", "WRITE_SIZE", 10.0),
                          ("wf_engine_async<2, 4, 32>", "SQ_INSTS_VALU_INT32", 1000)])
    assert P.split_widths(1025) == [512, 513, 513]
    load_b, store_b = P.split_bytes(1025, 1024, 512)
    assert store_b == (512 + 513 + 513) * 8 + 2 * 1024 * 4
    assert load_b == sum(w * 8 + 2 * w + 1024 for w in (512, 513, 513))
    out = tmp_path / "m.md"
    P.main(["metrics", "--trace", str(tmp_path / "t"), "--pmc", str(tmp_path / "c"), "--kernel", "wf_engine_async",
            "--problem", "SW", "--n1", "1025", "--n2", "1024", "--strip-rows", "512", "--column-split",
            "--out", str(out)])
    text = out.read_text()
    assert f"{1000 * 64 / (1538 * 1024):.2f} (algorithmic: 8)" in text
    assert "pairs of [512, 513, 513] columns, 1574912 cells" in text
