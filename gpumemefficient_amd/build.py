"""Build every native artefact of the package in-tree.

    python -m gpumemefficient_amd.build            # library + drivers + oracle
    python -m gpumemefficient_amd.build --lib      # libwfdp.so only

Outputs (git-ignored, shipped with the source tree):
    gpumemefficient_amd/lib/libwfdp.so          C ABI of include/wfdp.h (HIP, gfx950)
    gpumemefficient_amd/lib/libwfdp_compat_share.so, libwfdp_compat_cache.so
                                                the reference's C++ signatures
    gpumemefficient_amd/lib/libwfdp_compat_sat.so, libwfdp_compat_sor.so
                                                SOR() of the SAT and 2D-SOR programs
    gpumemefficient_amd/bin/<Problem>/<driver>  drop-in drivers (LCS, SmithWaterman,
                                                EditDistance, LCS-SHFL, SAT, 2D-SOR)
    oracle/liboracle.so, oracle/_ref/*          checker side (oracle/Makefile)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_obj"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WFDP_ARCH", "gfx950")
# Packed fp32 VALU forms are disabled for all device code.
HIP_FLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
CXX = os.environ.get("CXX", "g++")

# Drivers: (problem directory, binary name, problem, style)
#   style "share5": argv e1 e2 tX tY itr   (SmithWaterman/EditDistance Share)
#   style "cache2": argv e1 e2, itr = 100  (Cache drivers, LCS Share)
#   style "shfl":   argv e1 e2, prints total seconds of 100 calls
#   style "naive", "tile": argv e1 e2, one call, prints its seconds
DRIVERS = [
    ("LCS", "GPU_Hyperlane_Share", "LCS", "lcs2", "share"),
    ("LCS", "GPU_Hyperlane_Cache", "LCS", "lcs2", "cache"),
    ("SmithWaterman", "GPU_Hyperlane_Share", "SW", "share5", "share"),
    ("SmithWaterman", "GPU_Hyperlane_Cache", "SW", "cache2", "cache"),
    ("EditDistance", "GPU_Hyperlane_Share", "ED", "share5", "share"),
    ("EditDistance", "GPU_Hyperlane_Cache", "ED", "cache2", "cache"),
    ("LCS-SHFL", "GPU_Hyperlane_Share", "LCS", "shfl", "share"),
    ("LCS-SHFL", "GPU_Hyperlane_Cache", "LCS", "shfl", "cache"),
    # Naive baselines: GPU (one launch per cell anti-diagonal) and
    # GPU_TILE_SHARE (one launch per tile anti-diagonal; the reference's
    # SmithWaterman and EditDistance copies of it compute LCS).
    ("LCS", "GPU", "LCS", "naive", "naive"),
    ("SmithWaterman", "GPU", "SW", "naive", "naive"),
    ("EditDistance", "GPU", "ED", "naive", "naive"),
    ("LCS", "GPU_TILE_SHARE", "LCS", "tile", "tile"),
    ("SmithWaterman", "GPU_TILE_SHARE", "LCS", "tile", "tile"),
    ("EditDistance", "GPU_TILE_SHARE", "LCS", "tile", "tile"),
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _hip_obj(src: Path, obj: Path) -> None:
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])


def build_lib() -> Path:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / "wf_engine.hip", CSRC / "wf_kernels.hip", CSRC / "wf_api.cpp", CSRC / "wf_server.cpp",
            CSRC / "wf_tune.cpp", CSRC / "wf_naive.hip", CSRC / "wf_stencil.hip", CSRC / "wf_stencil_api.cpp", CSRC / "wf_bitpar.hip",
            CSRC / "wf_noalloc.hip"]
    srcs = [s for s in srcs if s.exists()]
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        futs = [ex.submit(_hip_obj, s, o) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    out = LIB / "libwfdp.so"
    tmp = LIB / "libwfdp.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_compat() -> None:
    src = CSRC / "wf_compat.cpp"
    if not src.exists():
        return
    for policy in ("share", "cache", "naive", "tile", "sat", "sor"):
        out = LIB / f"libwfdp_compat_{policy}.so"
        _run([CXX, "-std=c++17", "-O2", "-fPIC", "-shared", f"-DWF_COMPAT_POLICY_{policy.upper()}",
              f"-I{INCLUDE}", str(src), "-o", str(out), f"-L{LIB}", "-lwfdp",
              "-Wl,-rpath,$ORIGIN"])


def build_drivers() -> None:
    src = CSRC / "wf_driver.cpp"
    if not src.exists():
        return
    for pdir, name, prob, style, policy in DRIVERS:
        d = BIN / pdir
        d.mkdir(parents=True, exist_ok=True)
        _run([CXX, "-std=c++17", "-O2", f"-DWF_DRIVER_PROBLEM_{prob}", f"-DWF_DRIVER_STYLE_{style.upper()}",
              f"-I{INCLUDE}", str(src), "-o", str(d / name),
              f"-L{LIB}", f"-lwfdp_compat_{policy}", "-lwfdp", "-Wl,-rpath,$ORIGIN/../../lib"])
    gen = CSRC / "wf_generator.cpp"
    if gen.exists():
        for pdir in ("LCS", "SmithWaterman", "EditDistance", "LCS-SHFL"):
            d = BIN / pdir
            d.mkdir(parents=True, exist_ok=True)
            minexp = 4 if pdir in ("LCS", "SmithWaterman") else 8
            _run([CXX, "-std=c++17", "-O2", f"-DWF_GEN_MIN_EXP={minexp}", str(gen), "-o", str(d / "generator")])


# Stencil drivers: (program directory, binary name, compat library, style)
STENCIL_DRIVERS = [
    ("SAT", "GPU_Hyperlane_Share", "sat", "share5"),
    ("SAT", "GPU_Hyperlane_Cache", "sat", "cache2"),
    ("2D-SOR", "GPU_Hyperlane_Share", "sor", "share5"),
    ("2D-SOR", "GPU_Hyperlane_Cache", "sor", "cache2"),
]


def build_stencil_drivers(min_exp: int = 13) -> None:
    src = CSRC / "wf_stencil_driver.cpp"
    for pdir, name, lib, style in STENCIL_DRIVERS:
        d = BIN / pdir
        d.mkdir(parents=True, exist_ok=True)
        flags = ["-DWF_SDRIVER_SHARE5"] if style == "share5" else []
        _run([CXX, "-std=c++17", "-O2", *flags, f'-DWF_SDRIVER_NAME="{pdir}"', f"-I{INCLUDE}", str(src),
              "-o", str(d / name), f"-L{LIB}", f"-lwfdp_compat_{lib}", "-lwfdp", "-Wl,-rpath,$ORIGIN/../../lib"])
    gen = CSRC / "wf_stencil_generator.cpp"
    for pdir in ("SAT", "2D-SOR"):
        _run([CXX, "-std=c++17", "-O2", f"-DWF_GEN_MIN_EXP={min_exp}", str(gen), "-o", str(BIN / pdir / "generator")])


def build_oracle() -> None:
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    _run(["make", "-s", "-C", str(ROOT / "oracle")])


def build_all(lib_only: bool = False) -> None:
    build_lib()
    if lib_only:
        return
    build_compat()
    build_drivers()
    build_stencil_drivers()
    build_oracle()


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lib", action="store_true", help="build libwfdp.so only")
    args = ap.parse_args(argv)
    build_all(lib_only=args.lib)
    print("built", LIB / "libwfdp.so")
    return 0


if __name__ == "__main__":
    sys.exit(main())
