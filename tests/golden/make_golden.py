"""Writes tests/golden/ref_vectors.json: inputs and the answers printed by the
reference's own CPU programs (LCS/LCS.cpp, SmithWaterman/SW.cpp,
EditDistance/ED.cpp), compiled from source by oracle/Makefile into
oracle/_ref/.

Run in the build container (needs /root/reference):
    make -C oracle ref && python tests/golden/make_golden.py

The reference programs read ./Data/str1_2_<e1>_str2_2_<e2>.txt and are only
defined for n1 == n2 (LCS/LCS.cpp:54-59), so every case here is square.  The
file name only selects the file; the sizes come from its header, so any n
can be run as "str1_2_1_str2_2_1.txt".
"""
from __future__ import annotations

import hashlib
import json
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
from gpumemefficient_amd.datafile import format_input, read_input  # noqa: E402

REF_BIN = ROOT / "oracle" / "_ref"
REF_TREE = Path("/root/reference")


def run_ref(prog: str, a: np.ndarray, b: np.ndarray) -> int:
    with tempfile.TemporaryDirectory() as td:
        (Path(td) / "Data").mkdir()
        (Path(td) / "Data" / "str1_2_1_str2_2_1.txt").write_text(format_input(a, b))
        out = subprocess.run([str(REF_BIN / prog), "1", "1"], cwd=td, capture_output=True, text=True,
                             check=True).stdout
    m = re.search(r"(?:lcs length|last): (-?\d+)", out)
    if not m:
        raise RuntimeError(f"unexpected output from {prog}: {out!r}")
    return int(m.group(1))


def main() -> int:
    for p in ("LCS", "SW", "ED"):
        if not (REF_BIN / p).exists():
            print(f"missing {REF_BIN / p}: run `make -C oracle ref` first", file=sys.stderr)
            return 1
    rng = np.random.default_rng(20260)
    cases = []
    sizes = [1, 2, 3, 4, 5, 7, 8, 13, 16, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 500, 1000]
    for n in sizes:
        for alpha in (2, 4, 20):
            a = rng.integers(0, alpha, n, dtype=np.int32)
            b = rng.integers(0, alpha, n, dtype=np.int32)
            cases.append({"n": n, "alpha": alpha, "a": a.tolist(), "b": b.tolist()})
    # Identical strings, one planted homologous block, and the ED row-0 quirk.
    a = rng.integers(0, 4, 300, dtype=np.int32)
    cases.append({"n": 300, "alpha": 4, "a": a.tolist(), "b": a.tolist(), "note": "identical"})
    b = rng.integers(0, 4, 300, dtype=np.int32)
    b[100:200] = a[50:150]
    cases.append({"n": 300, "alpha": 4, "a": a.tolist(), "b": b.tolist(), "note": "planted block"})
    cases.append({"n": 2, "alpha": 4, "a": [1, 2], "b": [2, 3], "note": "ED.cpp row-0 quirk (textbook 2)"})
    cases.append({"n": 2, "alpha": 4, "a": [0, 1], "b": [1, 0], "note": "ED.cpp row-0 quirk (textbook 2)"})
    for c in cases:
        a, b = np.array(c["a"], np.int32), np.array(c["b"], np.int32)
        c["lcs"] = run_ref("LCS", a, b)
        c["sw_last"] = run_ref("SW", a, b)
        c["ed"] = run_ref("ED", a, b)

    # The reference's committed square inputs (LCS/Data, byte-identical in
    # LCS-SHFL/Data).  Small ones are copied to tests/golden/data/.
    ref_files = {}
    data_dir = REF_TREE / "LCS" / "Data"
    if data_dir.exists():
        for e in range(8, 16):
            f = data_dir / f"str1_2_{e}_str2_2_{e}.txt"
            a, b = read_input(f)
            ref_files[f.name] = {
                "md5": hashlib.md5(f.read_bytes()).hexdigest(),
                "lcs": run_ref("LCS", a, b), "sw_last": run_ref("SW", a, b), "ed": run_ref("ED", a, b),
            }
            if e in (8, 10, 12, 15):
                (HERE / "data").mkdir(exist_ok=True)
                (HERE / "data" / f.name).write_bytes(f.read_bytes())
    out = {"generated_by": "tests/golden/make_golden.py", "programs": "oracle/_ref (LCS.cpp, SW.cpp, ED.cpp)",
           "published": {"str1_2_15_str2_2_15.txt": {"lcs": 21469,
                                                     "source": "LCS/GPU_Hyperlane_Share.txt:102"}},
           "cases": cases, "reference_data": ref_files}
    (HERE / "ref_vectors.json").write_text(json.dumps(out, separators=(",", ":")))
    print(f"wrote {len(cases)} cases, {len(ref_files)} reference data files")
    return 0


if __name__ == "__main__":
    sys.exit(main())
