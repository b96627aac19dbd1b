"""The reference's CPU programs as drop-ins (oracle/bin/{LCS,SW,ED}, built
from oracle/wf_cpu_main.c over the C restatement) and the all-cores
anti-diagonal variant of the restatement.

The programs must print what LCS/LCS.cpp:137-138, SmithWaterman/SW.cpp
:136-139 and EditDistance/ED.cpp:134-137 print, with the scores the
reference programs print on the reference's own data files
(tests/golden/ref_vectors.json, written by the reference programs compiled
from source, oracle/_ref), in all three modes: rolling rows, --faithful (the
whole table, as the reference allocates it) and --threads."""
import json
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "oracle" / "bin"
GOLD = ROOT / "tests" / "golden"
VEC = json.loads((GOLD / "ref_vectors.json").read_text())["reference_data"]
REF_DATA = Path("/root/reference/LCS/Data")
PROGRAMS = {"LCS": ("lcs", "lcs length: ", " ms."), "SW": ("sw_last", "last: ", " second."),
            "ED": ("ed", "last: ", " ms.")}


@pytest.fixture(scope="module", autouse=True)
def built():
    O.build()


def _data_dir(tmp_path: Path, e: int) -> Path:
    name = f"str1_2_{e}_str2_2_{e}.txt"
    src = GOLD / "data" / name
    if not src.exists():
        src = REF_DATA / name
    if not src.exists():
        pytest.skip(f"{name} is neither committed nor in a mounted reference tree")
    (tmp_path / "Data").mkdir(exist_ok=True)
    shutil.copy(src, tmp_path / "Data" / name)
    return tmp_path


def _run(prog: str, cwd: Path, *args: str) -> subprocess.CompletedProcess:
    return subprocess.run([str(BIN / prog), *args], cwd=cwd, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("prog", sorted(PROGRAMS))
def test_argv_and_missing_file(tmp_path, prog):
    r = _run(prog, tmp_path)
    assert r.returncode == 1 and r.stdout == "Incorrect Input Parameters. Must be two string sizes.\n"
    r = _run(prog, tmp_path, "7", "7")
    assert r.returncode == 1 and r.stdout == "ERROR: Input file cannot be opened!\n"


@pytest.mark.parametrize("e", [8, 9, 10, 11, 12, 13])
@pytest.mark.parametrize("prog", sorted(PROGRAMS))
@pytest.mark.parametrize("mode", [[], ["--faithful"], ["--threads", "5"]], ids=["rows", "faithful", "threads"])
def test_reference_outputs(tmp_path, prog, e, mode):
    key, head, tail = PROGRAMS[prog]
    cwd = _data_dir(tmp_path, e)
    r = _run(prog, cwd, str(e), str(e), *mode)
    assert r.returncode == 0, r.stdout
    lines = r.stdout.splitlines()
    want = VEC[f"str1_2_{e}_str2_2_{e}.txt"][key]
    assert lines[0] == f"{head}{want}"
    assert re.fullmatch(r"execution time: [0-9.e+-]+" + re.escape(tail), lines[1]), lines[1]
    assert len(lines) == (3 if mode[:1] == ["--threads"] else 2)


@pytest.mark.parametrize("mode", [["--faithful"], ["--threads", "8"]], ids=["faithful", "threads"])
def test_published_lcs_2_15(tmp_path, mode):
    """The one published hot-path result, lcs length 21469 at 2^15
    (LCS/GPU_Hyperlane_Share.txt:102); --faithful allocates the reference's
    4.3 GB table."""
    cwd = _data_dir(tmp_path, 15)
    r = _run("LCS", cwd, "15", "15", *mode)
    assert r.returncode == 0, r.stdout
    assert r.stdout.splitlines()[0] == "lcs length: 21469"


@pytest.mark.parametrize("prog", ["SW", "ED"])
def test_2_15_rows_and_threads(tmp_path, prog):
    key, head, _ = PROGRAMS[prog]
    cwd = _data_dir(tmp_path, 15)
    want = VEC["str1_2_15_str2_2_15.txt"][key]
    for mode in ([], ["--threads", "8"]):
        r = _run(prog, cwd, "15", "15", *mode)
        assert r.returncode == 0 and r.stdout.splitlines()[0] == f"{head}{want}"


@pytest.mark.parametrize("prob", ["LCS", "SW", "ED"])
def test_blocked_matches_rolling_rows(prob):
    """The anti-diagonal tile schedule against the rolling rows: non-square
    pairs, tile sizes that leave ragged last tiles, 1 to 9 threads, both ED
    boundaries.  A tile that read a stale corner or column would fail."""
    rng = np.random.default_rng(77)
    for n1, n2, alpha, tile, threads in [(1, 1, 2, 16, 1), (17, 300, 4, 16, 3), (300, 17, 4, 16, 9),
                                          (1000, 777, 4, 64, 4), (777, 1000, 20, 100, 5), (2049, 2047, 4, 256, 8),
                                          (5000, 300, 2, 33, 7)]:
        a = rng.integers(0, alpha, n1, dtype=np.int32)
        b = rng.integers(0, alpha, n2, dtype=np.int32)
        for textbook in ((False, True) if prob == "ED" else (False,)):
            assert O.blocked(prob, a, b, threads, tile, textbook) == O.score(prob, a, b, textbook), \
                (n1, n2, tile, threads, textbook)
