"""The reference's input-file format and file naming.

Format (LCS/generator.cpp:58-66): ``"n1 n2\\n"``, then the n1 symbols of arr1
each followed by a space, a newline, the n2 symbols of arr2 the same way, a
newline.  Readers accept any whitespace, like readInputData's ``>>``
(LCS/GPU_Hyperlane_Share.cpp:14-34).  Files are named
``./Data/str1_2_<e1>_str2_2_<e2>.txt`` (LCS/GPU_Hyperlane_Share.cpp:65-79).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def data_path(e1: int, e2: int, root: str | Path = ".") -> Path:
    return Path(root) / "Data" / f"str1_2_{e1}_str2_2_{e2}.txt"


def read_input(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    toks = Path(path).read_bytes().split()
    if len(toks) < 2:
        raise ValueError(f"{path}: missing header")
    n1, n2 = int(toks[0]), int(toks[1])
    vals = np.array(toks[2:2 + n1 + n2], dtype=np.int64)
    if vals.size != n1 + n2:
        raise ValueError(f"{path}: expected {n1 + n2} symbols, found {vals.size}")
    return vals[:n1].astype(np.int32), vals[n1:].astype(np.int32)


def format_input(a, b) -> str:
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return (f"{a.size} {b.size}\n" + "".join(f"{int(v)} " for v in a) + "\n"
            + "".join(f"{int(v)} " for v in b) + "\n")


def write_input(path: str | Path, a, b) -> Path:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(format_input(a, b))
    return p
