"""ctypes binding of oracle/liboracle.so -- the CPU restatement of
LCS/LCS.cpp, SmithWaterman/SW.cpp and EditDistance/ED.cpp.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product path.

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  lcs  golden vectors from the reference LCS.cpp compiled from source, the
       published 21469 at 2^15, brute force over subsequences, closed forms
  sw   golden vectors from SW.cpp, brute force over local alignments (max),
       closed forms
  ed   golden vectors from ED.cpp (including the row-0 quirk), brute force of
       the textbook Levenshtein distance in textbook mode, closed forms
  table / last_row  checked against lcs/sw/ed and the reference full tables
  blocked  (the all-cores anti-diagonal variant) checked against lcs/sw/ed
       on random non-square pairs, tiles that do not divide the sizes and
       thread counts from 1 to 9, and through the CPU drop-in programs
       (bin/LCS, bin/SW, bin/ED) against the reference's own outputs
  stencil  SAT: the closed form (2-D prefix sums of the interior plus the
       boundary row and column, mod 2^32); SOR: hand-computed known answers
       (which neighbours are new, which old; truncation of negative sums),
       constant tables as fixed points, and an anti-diagonal traversal that
       must give the row-major result
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

LCS, ED, SW = 0, 1, 2
_NAMES = {"LCS": LCS, "ED": ED, "SW": SW}
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "oracle", "cpu"], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    I32P = C.POINTER(C.c_int32)
    lib.oracle_lcs.argtypes = [I32P, C.c_int64, I32P, C.c_int64]
    lib.oracle_lcs.restype = C.c_int32
    lib.oracle_sw.argtypes = [I32P, C.c_int64, I32P, C.c_int64, I32P]
    lib.oracle_sw.restype = C.c_int32
    lib.oracle_ed.argtypes = [I32P, C.c_int64, I32P, C.c_int64, C.c_int]
    lib.oracle_ed.restype = C.c_int32
    lib.oracle_table.argtypes = [C.c_int, I32P, C.c_int64, I32P, C.c_int64, C.c_int, I32P]
    lib.oracle_table.restype = C.c_int32
    lib.oracle_last_row.argtypes = [C.c_int, I32P, C.c_int64, I32P, C.c_int64, C.c_int, I32P]
    lib.oracle_last_row.restype = C.c_int
    lib.oracle_stencil.argtypes = [C.c_int, C.c_int64, C.c_int64, I32P, C.c_int64]
    lib.oracle_stencil.restype = None
    lib.oracle_blocked.argtypes = [C.c_int, I32P, C.c_int64, I32P, C.c_int64, C.c_int, C.c_int, C.c_int64, I32P]
    lib.oracle_blocked.restype = C.c_int32
    _lib = lib
    return lib


def _a(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))


def _p(x: np.ndarray):
    return x.ctypes.data_as(C.POINTER(C.c_int32))


def problem_id(p) -> int:
    return _NAMES[p.upper()] if isinstance(p, str) else int(p)


def lcs(a, b) -> int:
    a, b = _a(a), _a(b)
    return int(load().oracle_lcs(_p(a), a.size, _p(b), b.size))


def sw(a, b) -> tuple[int, int]:
    """Returns (last, max)."""
    a, b = _a(a), _a(b)
    m = C.c_int32(0)
    last = load().oracle_sw(_p(a), a.size, _p(b), b.size, C.byref(m))
    return int(last), int(m.value)


def ed(a, b, textbook: bool = False) -> int:
    a, b = _a(a), _a(b)
    return int(load().oracle_ed(_p(a), a.size, _p(b), b.size, 1 if textbook else 0))


def score(problem, a, b, textbook: bool = False) -> tuple[int, int]:
    """(last, max) for any problem; max is 0 except for SW."""
    p = problem_id(problem)
    if p == LCS:
        return lcs(a, b), 0
    if p == SW:
        return sw(a, b)
    return ed(a, b, textbook), 0


def blocked(problem, a, b, threads: int, tile: int = 256, textbook: bool = False) -> tuple[int, int]:
    """(last, max) computed by `threads` threads over anti-diagonals of
    tile x tile blocks."""
    a, b = _a(a), _a(b)
    m = C.c_int32(0)
    last = load().oracle_blocked(problem_id(problem), _p(a), a.size, _p(b), b.size, 1 if textbook else 0,
                                 int(threads), int(tile), C.byref(m))
    return int(last), int(m.value)


def table(problem, a, b, textbook: bool = False) -> np.ndarray:
    a, b = _a(a), _a(b)
    t = np.zeros((b.size + 1, a.size + 1), dtype=np.int32)
    load().oracle_table(problem_id(problem), _p(a), a.size, _p(b), b.size, 1 if textbook else 0, _p(t))
    return t


def last_row(problem, a, b, textbook: bool = False) -> np.ndarray:
    a, b = _a(a), _a(b)
    out = np.zeros(a.size + 1, dtype=np.int32)
    load().oracle_last_row(problem_id(problem), _p(a), a.size, _p(b), b.size, 1 if textbook else 0, _p(out))
    return out


SAT, SOR = 0, 1


def stencil(op, table: np.ndarray) -> np.ndarray:
    """One in-place sweep (op "SAT"/"SOR" or 0/1) of an (n2+2) x ld int32
    table with ld >= n1+3; returns the updated copy.  n1 = ld - 3 unless
    the table carries more columns: pass a view of shape (n2+2, n1+3)."""
    op = {"SAT": SAT, "SOR": SOR}[op.upper()] if isinstance(op, str) else int(op)
    t = np.array(table, dtype=np.int32, copy=True, order="C")
    n2 = t.shape[0] - 2
    n1 = t.shape[1] - 3
    if n1 < 0 or n2 < 0:
        raise ValueError("table must have at least 2 rows and 3 columns")
    load().oracle_stencil(op, n1, n2, _p(t), t.shape[1])
    return t
