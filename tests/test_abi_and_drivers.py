"""C ABI exports, drop-in driver argv/stdout contract, data-file format.

The CPU tests here make no compute call; the GPU tests run each drop-in driver
binary on the reference's committed inputs (tests/golden/data) and compare
its stdout with the reference driver format (LCS/GPU_Hyperlane_Share.cpp
:98-108, SmithWaterman/GPU_Hyperlane_Share.cpp:62,93,107-112,
EditDistance/GPU_Hyperlane_Share.cpp:91,111-115, the Cache drivers and
LCS-SHFL/GPU_Hyperlane_Share.cpp:98-107)."""
import ctypes
import json
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpumemefficient_amd import _lib
from gpumemefficient_amd.datafile import format_input, read_input, write_input

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "gpumemefficient_amd" / "bin"
LIBDIR = ROOT / "gpumemefficient_amd" / "lib"
GOLD = ROOT / "tests" / "golden"
VEC = json.loads((GOLD / "ref_vectors.json").read_text())

DRIVERS = [
    ("LCS", "GPU_Hyperlane_Share"), ("LCS", "GPU_Hyperlane_Cache"),
    ("SmithWaterman", "GPU_Hyperlane_Share"), ("SmithWaterman", "GPU_Hyperlane_Cache"),
    ("EditDistance", "GPU_Hyperlane_Share"), ("EditDistance", "GPU_Hyperlane_Cache"),
    ("LCS-SHFL", "GPU_Hyperlane_Share"), ("LCS-SHFL", "GPU_Hyperlane_Cache"),
    ("LCS", "GPU"), ("SmithWaterman", "GPU"), ("EditDistance", "GPU"),
    ("LCS", "GPU_TILE_SHARE"), ("SmithWaterman", "GPU_TILE_SHARE"), ("EditDistance", "GPU_TILE_SHARE"),
    ("SAT", "GPU_Hyperlane_Share"), ("SAT", "GPU_Hyperlane_Cache"),
    ("2D-SOR", "GPU_Hyperlane_Share"), ("2D-SOR", "GPU_Hyperlane_Cache"),
]
HYPERLANE = [d for d in DRIVERS if d[1].startswith("GPU_Hyperlane") and d[0] not in ("SAT", "2D-SOR")]


def _header_functions(path: Path) -> set[str]:
    txt = re.sub(r"/\*.*?\*/", "", path.read_text(), flags=re.S)
    return set(re.findall(r"\b(wf_[a-z_0-9]+)\s*\(", txt))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _header_functions(ROOT / "include" / "wfdp.h")
    assert declared, "no declarations parsed"
    for name in sorted(declared):
        assert hasattr(lib, name), f"libwfdp.so does not export {name}"
    assert set(_lib.EXPORTS) <= declared
    assert lib.wf_abi_version() == 1
    assert lib.wf_strerror(_lib.WF_ERR_TIMEOUT).decode().startswith("device-side")


def test_compat_libraries_export_reference_signatures():
    # Mangled C++ names of the reference's six host functions (and LCS-SHFL's).
    want = ["_Z3LCSiiPiS_iiS_", "_Z3LCSiiPiS_", "_Z2SWiiPiS_iiS_ii", "_Z2SWiiPiS_iiS_",
            "_Z2EDiiPiS_iiS_ii", "_Z2EDiiPiS_iiS_"]
    for policy in ("share", "cache"):
        p = LIBDIR / f"libwfdp_compat_{policy}.so"
        out = subprocess.run(["nm", "-D", "--defined-only", str(p)], capture_output=True, text=True,
                             check=True).stdout
        for sym in want:
            assert sym in out, f"{p.name} lacks {sym}"
    # The naive baselines' four-argument functions (GPU.h, GPU_TILE_SHARE.h).
    for policy, syms in (("naive", ["_Z3LCSiiPiS_", "_Z2SWiiPiS_", "_Z2EDiiPiS_"]), ("tile", ["_Z3LCSiiPiS_"])):
        p = LIBDIR / f"libwfdp_compat_{policy}.so"
        out = subprocess.run(["nm", "-D", "--defined-only", str(p)], capture_output=True, text=True,
                             check=True).stdout
        for sym in syms:
            assert sym in out, f"{p.name} lacks {sym}"
    # SOR(int, int, int*, int, int) and SOR(int, int, int*) of the SAT and
    # 2D-SOR programs (SAT/GPU_Hyperlane_Share.h:5, SAT/GPU_Hyperlane_Cache.h:5).
    for policy in ("sat", "sor"):
        p = LIBDIR / f"libwfdp_compat_{policy}.so"
        out = subprocess.run(["nm", "-D", "--defined-only", str(p)], capture_output=True, text=True,
                             check=True).stdout
        for sym in ("_Z3SORiiPiii", "_Z3SORiiPi"):
            assert sym in out, f"{p.name} lacks {sym}"


def test_invalid_arguments_rejected_without_device():
    lib = _lib.load()
    out = _lib.wf_result()
    # Null context: rejected before any HIP call.
    assert lib.wf_align(None, 0, None, 0, None, 0, None, ctypes.byref(out)) == _lib.WF_ERR_INVALID
    assert lib.wf_plan_run(None, 0) == _lib.WF_ERR_INVALID
    assert lib.wf_ctx_set_timeout(None, 1.0) == _lib.WF_ERR_INVALID
    assert lib.wf_stencil(None, 0, 1, 1, None, 4, None, ctypes.byref(out)) == _lib.WF_ERR_INVALID
    assert lib.wf_stencil_device(None, 1, 1, 1, None, 4, None, ctypes.byref(out)) == _lib.WF_ERR_INVALID


# The usage line each reference driver prints for a wrong argc, verbatim.
USAGE_TWO = "Incorrect Input Parameters. Must be two string sizes."
USAGE = {
    # SmithWaterman/GPU_Hyperlane_Share.cpp:51 (its own spacing)
    ("SmithWaterman", "GPU_Hyperlane_Share"):
        "Incorrect Input Parameters. Must be two string sizes ,two tile sizes and one total iteration step.",
    # EditDistance/GPU_Hyperlane_Share.cpp:51
    ("EditDistance", "GPU_Hyperlane_Share"):
        "Incorrect Input Parameters. Must be two string sizes, two tile sizes, and one total iteration step.",
    # SAT/GPU_Hyperlane_Share.cpp:46
    ("SAT", "GPU_Hyperlane_Share"):
        "Incorrect Input Parameters. Must be two string sizes, two tile sizes, and one iteration step.",
    # 2D-SOR/nested_loop/GPU_Hyperlane_Share.cpp:46
    ("2D-SOR", "GPU_Hyperlane_Share"):
        "Incorrect Input Parameters. Must be two string sizes, two tile sizes, and one total iteration step.",
}


@pytest.mark.parametrize("pdir,name", DRIVERS)
def test_driver_argv_errors(tmp_path, pdir, name):
    """Wrong argc and a missing input file: exit code 1 and the reference's
    lines, whole (LCS/GPU_Hyperlane_Share.cpp:52, :19 and the other drivers'
    counterparts)."""
    exe = BIN / pdir / name
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stdout == USAGE.get((pdir, name), USAGE_TWO) + "\n"
    args = ["9", "9", "32", "128", "1"] if (pdir, name) in USAGE else ["9", "9"]
    r = subprocess.run([str(exe), *args], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    lines = r.stdout.splitlines()
    assert lines[-1] == "ERROR: Input file cannot be opened!", lines


REF = Path("/root/reference")
REF_DRIVERS = [(d, f) for d in ("LCS", "SmithWaterman", "EditDistance", "LCS-SHFL") for f in ("Share", "Cache")]


@pytest.mark.skipif(not REF.is_dir(), reason="the reference tree is not mounted on this host")
@pytest.mark.parametrize("pdir,kind", REF_DRIVERS)
def test_reference_driver_sources_link_against_compat_libs(tmp_path, pdir, kind):
    """The reference's own driver sources, compiled where they lie and linked
    against libwfdp_compat_{share,cache}.so instead of the reference's .cu
    objects.  The SmithWaterman and EditDistance Cache drivers include the
    Share header (GPU_Hyperlane_Cache.cpp:7) but call the 7-argument form
    (:101), so every driver is compiled with wfdp_compat.h force-included,
    which declares both forms.  Link only: running needs a GPU."""
    src = REF / pdir / f"GPU_Hyperlane_{kind}.cpp"
    exe = tmp_path / f"{pdir}_{kind}"
    cmd = ["g++", "-std=c++11", "-O1", "-w", "-include", str(ROOT / "include" / "wfdp_compat.h"),
           f"-I{REF / pdir}", str(src), "-o", str(exe), f"-L{LIBDIR}", f"-lwfdp_compat_{kind.lower()}", "-lwfdp",
           f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    undefined = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    host_fn = {"LCS": "_Z3LCSiiPiS_iiS_", "LCS-SHFL": "_Z3LCSiiPiS_iiS_",
               "SmithWaterman": "_Z2SWiiPiS_iiS_ii" if kind == "Share" else "_Z2SWiiPiS_iiS_",
               "EditDistance": "_Z2EDiiPiS_iiS_ii" if kind == "Share" else "_Z2EDiiPiS_iiS_"}[pdir]
    assert host_fn in undefined
    exported = subprocess.run(["nm", "-D", "--defined-only", str(LIBDIR / f"libwfdp_compat_{kind.lower()}.so")],
                              capture_output=True, text=True, check=True).stdout
    assert host_fn in exported


def test_datafile_roundtrip(tmp_path):
    a = np.array([0, 3, 1, 2], np.int32)
    b = np.array([2, 2], np.int32)
    txt = format_input(a, b)
    assert txt == "4 2\n0 3 1 2 \n2 2 \n"  # LCS/generator.cpp:58-66 layout
    p = write_input(tmp_path / "Data" / "x.txt", a, b)
    a2, b2 = read_input(p)
    assert a2.tolist() == a.tolist() and b2.tolist() == b.tolist()
    for f in (GOLD / "data").glob("*.txt"):
        a, b = read_input(f)
        assert f.read_text() == format_input(a, b)


def test_generator_prefix_family(tmp_path):
    exe = BIN / "LCS" / "generator"
    subprocess.run([str(exe), "7", "7", "123"], cwd=tmp_path, check=True)
    files = sorted((tmp_path / "Data").glob("*.txt"))
    assert [f.name for f in files] == sorted(f"str1_2_{e}_str2_2_{e}.txt" for e in range(4, 8))
    a7, b7 = read_input(tmp_path / "Data" / "str1_2_7_str2_2_7.txt")
    a5, b5 = read_input(tmp_path / "Data" / "str1_2_5_str2_2_5.txt")
    assert a7.size == 128 and a5.tolist() == a7[:32].tolist() and b5.tolist() == b7[:32].tolist()
    assert set(a7.tolist()) <= {0, 1, 2, 3}
    # Same seed, same files.
    other = tmp_path / "again"
    other.mkdir()
    subprocess.run([str(exe), "7", "7", "123"], cwd=other, check=True)
    assert (other / "Data" / "str1_2_7_str2_2_7.txt").read_bytes() == (tmp_path / "Data" / "str1_2_7_str2_2_7.txt").read_bytes()


# ------------------------------------------------------------------ GPU tests
def _stage(tmp_path: Path, e: int) -> None:
    (tmp_path / "Data").mkdir(exist_ok=True)
    shutil.copy(GOLD / "data" / f"str1_2_{e}_str2_2_{e}.txt", tmp_path / "Data")


@pytest.mark.gpu
@pytest.mark.parametrize("pdir,name", HYPERLANE)
@pytest.mark.parametrize("e", [8, 10, 12])
def test_driver_stdout_and_score(tmp_path, pdir, name, e):
    _stage(tmp_path, e)
    rec = VEC["reference_data"][f"str1_2_{e}_str2_2_{e}.txt"]
    share5 = name.endswith("Share") and pdir in ("SmithWaterman", "EditDistance")
    args = [str(e), str(e)] + (["32", "128", "3"] if share5 else [])
    r = subprocess.run([str(BIN / pdir / name), *args], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    if pdir == "SmithWaterman" and share5:
        assert lines[0] == f"{e} {e} 32 128"
        lines = lines[1:]
    assert lines[0] == "input data loaded"
    if pdir in ("LCS", "LCS-SHFL"):
        assert lines[1] == f"lcs length: {rec['lcs']}"
    elif pdir == "SmithWaterman":
        assert lines[1] == f"last: {rec['sw_last']}"
    else:
        assert lines[1] == f"last: {rec['ed']}"
    unit = "second." if pdir == "LCS-SHFL" else "ms."
    assert re.fullmatch(r"execution time: [0-9.e+-]+ " + re.escape(unit), lines[2]), lines[2]
    assert len(lines) == 3


@pytest.mark.gpu
def test_lcs_driver_published_2_15(tmp_path):
    """LCS/GPU_Hyperlane_Share.txt:102 prints `lcs length: 21469` on the
    committed 2^15 input; the drop-in prints the same (5-arg form, tile 256
    wide, 1024 rows, one call)."""
    _stage(tmp_path, 15)
    r = subprocess.run([str(BIN / "LCS" / "GPU_Hyperlane_Share"), "15", "15", "256", "1024", "1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert r.stdout.splitlines()[1] == "lcs length: 21469"


@pytest.mark.parametrize("prog", ["SAT", "2D-SOR"])
def test_stencil_generator_format(tmp_path, prog):
    """generator <e1> <e2> [seed] writes the reference's x_2_<e>_y_2_<f>.txt
    family (SAT/generator.cpp:43-66): "n1 n2", then n2 + 3 rows of n1 + 3
    values in [0, 10000), each smaller file the top-left corner of the larger."""
    exe = BIN / prog / "generator"
    r = subprocess.run([str(exe), "14", "13", "7"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted((tmp_path / "Data").iterdir())
    assert [f.name for f in files] == ["x_2_13_y_2_12.txt", "x_2_14_y_2_13.txt"]
    big = (tmp_path / "Data" / "x_2_14_y_2_13.txt").read_text().split("\n")
    small = (tmp_path / "Data" / "x_2_13_y_2_12.txt").read_text().split("\n")
    assert big[0] == "16384 8192" and small[0] == "8192 4096"
    rows = [ln for ln in big[1:] if ln]
    assert len(rows) == 8192 + 3
    first = [int(v) for v in rows[0].split()]
    assert len(first) == 16384 + 3 and all(0 <= v < 10000 for v in first)
    assert rows[0].endswith(" ")
    assert [int(v) for v in small[1].split()] == first[: 8192 + 3]
    r2 = subprocess.run([str(exe), "13", "13", "7"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r2.returncode == 0
