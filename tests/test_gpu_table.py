"""Cell-level parity: the whole DP table from the table-dump kernel
(wf_align_table) against the oracle's table, and the drivers' DEBUG-style dump
file (LCS/GPU_Hyperlane_Share.cpp:109-131) against the same.  Integer work,
so every cell must match."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd.datafile import read_input

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "gpumemefficient_amd" / "bin"
GOLD = Path(__file__).resolve().parent / "golden"

# (n1, n2, alphabet): one cell, sub-wave, a single ragged strip, several
# strips with ragged edges both ways, and an alphabet past the row-table
# limit (compare mode).
SHAPES = [(1, 1, 4), (5, 70, 4), (63, 65, 20), (200, 130, 4), (1000, 777, 20), (513, 2049, 4), (700, 1500, 200)]


def first_diff(got, want):
    bad = np.argwhere(got != want)
    if bad.size == 0:
        return None
    y, x = bad[0]
    return f"first bad cell H({y + 1}, {x + 1}): got {got[y, x]} want {want[y, x]} ({len(bad)} cells differ)"


@pytest.mark.parametrize("prob,textbook", [("LCS", False), ("ED", False), ("ED", True), ("SW", False)])
@pytest.mark.parametrize("n1,n2,alpha", SHAPES)
def test_table_matches_oracle(ctx, prob, textbook, n1, n2, alpha):
    rng = np.random.default_rng(n1 * 7919 + n2 * 31 + alpha)
    a = rng.integers(0, alpha, n1, dtype=np.int32)
    b = rng.integers(0, alpha, n2, dtype=np.int32)
    print(f"  table {prob} textbook={textbook} n1={n1} n2={n2} alpha={alpha}", flush=True)
    r, t = ctx.align_table(prob, a, b, ed_textbook=textbook)
    want = O.table(prob, a, b, textbook=textbook)[1:, 1:]
    assert r.status == 0
    assert t.shape == (n2, n1)
    assert first_diff(t, want) is None, first_diff(t, want)
    assert r.last == want[-1, -1]
    if prob == "SW":
        assert r.max_score == want.max(initial=0)


def test_table_matches_score_path(ctx):
    """The dump kernel and the score-only path agree on the score of a pair
    long enough for several strips and chunks."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 20, 3000, dtype=np.int32)
    b = rng.integers(0, 20, 2600, dtype=np.int32)
    for prob in ("LCS", "ED", "SW"):
        print(f"  table vs align {prob} 3000x2600", flush=True)
        r, t = ctx.align_table(prob, a, b)
        s = ctx.align(prob, a, b)
        assert (r.last, r.max_score) == (s.last, s.max_score)
        assert t[-1, -1] == s.last


@pytest.mark.parametrize("pdir,name,prob", [("LCS", "GPU_Hyperlane_Share", "LCS"),
                                            ("SmithWaterman", "GPU_Hyperlane_Cache", "SW"),
                                            ("EditDistance", "GPU_Hyperlane_Cache", "ED")])
def test_driver_debug_dump(tmp_path, pdir, name, prob):
    """WFDP_DEBUG_TABLE=1 makes the drop-in driver write
    ./Output/output_Hyperlane_<e>.txt like the reference's DEBUG build: first
    line n1, then n1 rows of n2 space-terminated values."""
    e = 8
    (tmp_path / "Data").mkdir()
    (tmp_path / "Output").mkdir()
    src = GOLD / "data" / f"str1_2_{e}_str2_2_{e}.txt"
    shutil.copy(src, tmp_path / "Data")
    a, b = read_input(src)
    env = dict(os.environ, WFDP_DEBUG_TABLE="1")
    print(f"  driver dump {pdir}/{name} e={e}", flush=True)
    r = subprocess.run([str(BIN / pdir / name), str(e), str(e)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    text = (tmp_path / "Output" / f"output_Hyperlane_{e}.txt").read_text()
    lines = text.split("\n")
    assert lines[0] == str(a.size)
    assert lines[-1] == ""
    rows = lines[1:-1]
    assert len(rows) == a.size
    assert all(row.endswith(" ") for row in rows)
    got = np.array([[int(v) for v in row.split()] for row in rows], dtype=np.int32)
    want = O.table(prob, a, b)[1:, 1:]
    assert first_diff(got, want) is None, first_diff(got, want)
