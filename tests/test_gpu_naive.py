"""The naive baselines (wf_align_naive; the reference's GPU and
GPU_TILE_SHARE programs): every cell of their HBM table against the oracle,
their scores against the strip engines, and the drop-in GPU and
GPU_TILE_SHARE drivers' stdout.  Integer work: bit-exact."""
import json
import re
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
VEC = json.loads((GOLD / "ref_vectors.json").read_text())

# (n1, n2, alphabet): one cell, a sliver, exact 64 x 128 tiles, ragged tiles
# both ways, and symbols far outside one byte (no alphabet limit here).
SHAPES = [(1, 1, 4), (5, 70, 4), (64, 128, 4), (100, 300, 20), (257, 129, 4), (1000, 777, 100000)]


@pytest.mark.parametrize("method", ["diagonal", "tile"])
@pytest.mark.parametrize("prob,textbook", [("LCS", False), ("ED", False), ("ED", True), ("SW", False)])
@pytest.mark.parametrize("n1,n2,alpha", SHAPES)
def test_naive_table_matches_oracle(ctx, method, prob, textbook, n1, n2, alpha):
    rng = np.random.default_rng(n1 * 131 + n2 * 7 + alpha)
    a = rng.integers(0, alpha, n1, dtype=np.int32)
    b = rng.integers(0, alpha, n2, dtype=np.int32)
    print(f"  naive {method} {prob} textbook={textbook} n1={n1} n2={n2} alpha={alpha}", flush=True)
    r, t = ctx.align_naive(prob, a, b, method=method, ed_textbook=textbook, want_table=True)
    want = O.table(prob, a, b, textbook=textbook)
    assert r.status == 0
    bad = np.argwhere(t != want[1:, 1:])
    assert bad.size == 0, f"first bad cell H{tuple(bad[0] + 1)}: {t[tuple(bad[0])]} vs {want[1:, 1:][tuple(bad[0])]}"
    assert r.last == want[-1, -1]
    if prob == "SW":
        assert r.max_score == want.max()


def test_naive_agrees_with_strip_engine(ctx):
    rng = np.random.default_rng(9)
    a = rng.integers(0, 4, 3000, dtype=np.int32)
    b = rng.integers(0, 4, 2500, dtype=np.int32)
    for prob in ("LCS", "ED", "SW"):
        s = ctx.align(prob, a, b)
        for method in ("diagonal", "tile"):
            print(f"  naive {method} vs engine {prob} 3000x2500", flush=True)
            r, _ = ctx.align_naive(prob, a, b, method=method)
            assert (r.last, r.max_score) == (s.last, s.max_score)


def test_naive_empty_sides(ctx):
    a = np.arange(5, dtype=np.int32)
    e = np.zeros(0, dtype=np.int32)
    for method in ("diagonal", "tile"):
        assert ctx.align_naive("ED", a, e, method=method)[0].last == 4
        assert ctx.align_naive("ED", a, e, method=method, ed_textbook=True)[0].last == 5
        assert ctx.align_naive("ED", e, a, method=method)[0].last == 5
        assert ctx.align_naive("SW", a, e, method=method)[0].last == 0


MEMLINE = re.compile(r"current GPU memory info FREE: \d+ Bytes, Total: \d+ Bytes\.colsize: (\d+), rowsize: (\d+), "
                     r"allocates: (\d+) Bytes\.")


@pytest.mark.parametrize("pdir,name", [("LCS", "GPU"), ("SmithWaterman", "GPU"), ("EditDistance", "GPU"),
                                       ("LCS", "GPU_TILE_SHARE"), ("SmithWaterman", "GPU_TILE_SHARE"),
                                       ("EditDistance", "GPU_TILE_SHARE")])
@pytest.mark.parametrize("e", [8, 10])
def test_naive_driver_stdout(tmp_path, pdir, name, e):
    """Lines of the reference's GPU.cpp / GPU_TILE_SHARE.cpp: the data line,
    the memory line printed inside the host function, the score (the
    reference's SmithWaterman and EditDistance GPU_TILE_SHARE compute LCS;
    its naive ED starts row 0 at x), the time of the one call in seconds."""
    (tmp_path / "Data").mkdir()
    src = GOLD / "data" / f"str1_2_{e}_str2_2_{e}.txt"
    shutil.copy(src, tmp_path / "Data")
    a, b = read_input(src)
    rec = VEC["reference_data"][src.name]
    print(f"  driver {pdir}/{name} e={e}", flush=True)
    r = subprocess.run([str(BIN / pdir / name), str(e), str(e)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "input data loaded"
    m = MEMLINE.fullmatch(lines[1])
    assert m, lines[1]
    assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (a.size + 1, b.size + 1,
                                                                   (a.size + 1) * (b.size + 1) * 4)
    if name == "GPU_TILE_SHARE" or pdir == "LCS":
        assert lines[2] == f"lcs length: {rec['lcs']}"
    elif pdir == "SmithWaterman":
        assert lines[2] == f"last length: {rec['sw_last']}"
    else:
        assert lines[2] == f"last length: {O.score('ED', a, b, textbook=True)[0]}"
    assert re.fullmatch(r"execution time: [0-9.e+-]+ second\.", lines[3]), lines[3]
    assert len(lines) == 4
