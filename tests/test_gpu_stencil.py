"""The in-place table stencils (wf_stencil, wf_stencil_device; the reference's
SAT and 2D-SOR hyperlane programs): every cell of the swept table against the
oracle's row-major sweep, bit-exact, over ragged shapes, every strip height,
boundary-only tables, negative and wrapping values, the device-resident entry
point with a padded pitch (device memory from hipMalloc through ctypes), and
the drop-in drivers."""
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "gpumemefficient_amd" / "bin"

# (n1, n2): single cell, one row, one column, shorter than a wave's skew,
# exact 32-column blocks and 64-row waves, ragged everything, several strips
# of every height, more columns than the 128-column window, tall and thin.
SHAPES = [(1, 1), (9, 1), (1, 9), (50, 40), (64, 64), (96, 130), (128, 256), (257, 300), (1000, 517),
          (700, 1100), (3, 2000), (4097, 70)]


def _table(n1, n2, seed, lo=0, hi=10000):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, size=(n2 + 2, n1 + 3), dtype=np.int64).astype(np.int32)


def _check(got, want, what):
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{what}: {bad.shape[0]} cells differ, first {tuple(bad[0])}: {got[tuple(bad[0])]} vs " \
                          f"{want[tuple(bad[0])]}"


@pytest.mark.parametrize("op", ["SAT", "SOR"])
@pytest.mark.parametrize("n1,n2", SHAPES)
def test_stencil_matches_oracle(ctx, op, n1, n2):
    t = _table(n1, n2, n1 * 7919 + n2)
    print(f"  stencil {op} n1={n1} n2={n2}", flush=True)
    r, got = ctx.stencil(op, t)
    want = O.stencil(op, t)
    assert r.status == 0
    _check(got, want, f"{op} {n1}x{n2}")
    assert r.last == want[n2, n1]
    assert r.cells == n1 * n2


@pytest.mark.parametrize("strip_rows", [64, 128, 256])
@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_every_strip_height(ctx, op, strip_rows):
    n1, n2 = 333, 1000  # 16, 8 and 4 strips; the last one partial
    t = _table(n1, n2, strip_rows)
    print(f"  stencil {op} strip_rows={strip_rows}", flush=True)
    _, got = ctx.stencil(op, t, strip_rows=strip_rows)
    _check(got, O.stencil(op, t), f"{op} strips of {strip_rows}")


@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_signed_and_wrapping_values(ctx, op):
    # Full int32 range: SAT wraps mod 2^32, SOR's sums wrap and negative sums
    # truncate toward zero.
    n1, n2 = 200, 300
    t = _table(n1, n2, 5, lo=-(2 ** 31), hi=2 ** 31)
    print(f"  stencil {op} full-range values", flush=True)
    _, got = ctx.stencil(op, t)
    _check(got, O.stencil(op, t), f"{op} full range")


@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_boundary_only_tables(ctx, op):
    for n1, n2 in [(0, 0), (0, 5), (5, 0)]:
        t = _table(n1, n2, 3)
        r, got = ctx.stencil(op, t)
        assert r.status == 0
        _check(got, t, f"{op} {n1}x{n2} has no interior")


@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_device_resident_padded_pitch(ctx, op):
    from gpumemefficient_amd._hip import DeviceBuffer
    n1, n2 = 1021, 777
    ld = 1040  # >= n1 + 3, multiple of 4, padding columns beyond n1 + 2
    host = np.random.default_rng(11).integers(0, 10000, size=(n2 + 2, ld), dtype=np.int64).astype(np.int32)
    dev = DeviceBuffer.from_array(host)
    print(f"  stencil_device {op} n1={n1} n2={n2} ld={ld}", flush=True)
    r = ctx.stencil_device(op, dev.ptr, n1, n2, ld)
    got = dev.to_array()
    dev.free()
    want = host.copy()
    want[:, : n1 + 3] = O.stencil(op, host[:, : n1 + 3])
    _check(got, want, f"{op} device pitch {ld}")
    assert r.last == want[n2, n1]
    assert r.kernel_ms > 0


def test_stencil_device_rejects_bad_pitch(ctx):
    from gpumemefficient_amd import WfError
    from gpumemefficient_amd._hip import DeviceBuffer
    dev = DeviceBuffer((12, 16))
    with pytest.raises(WfError):
        ctx.stencil_device("SAT", dev.ptr, 10, 10, 14)  # pitch not a multiple of 4
    with pytest.raises(WfError):
        ctx.stencil_device("SAT", dev.ptr, 12, 10, 12)  # pitch < n1 + 3
    with pytest.raises(WfError):
        ctx.stencil_device("SAT", dev.ptr + 4, 8, 10, 12)  # base not 16-byte aligned
    dev.free()


@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_repeated_calls_are_independent(ctx, op):
    # The strip-boundary buffer and the launch epoch are shared with the
    # alignment engines; back-to-back launches of different shapes must not
    # see each other's boundary words.
    for n1, n2 in [(500, 900), (300, 1300), (500, 900)]:
        t = _table(n1, n2, n1 + n2)
        _, got = ctx.stencil(op, t)
        _check(got, O.stencil(op, t), f"{op} {n1}x{n2} repeated")
        a = np.arange(300, dtype=np.int32) % 4
        assert ctx.align("LCS", a, a).last == 300


@pytest.mark.parametrize("prog,op", [("SAT", "SAT"), ("2D-SOR", "SOR")])
def test_stencil_drivers(tmp_path, prog, op):
    """The drop-in drivers: argv and stdout of the reference programs, the
    generator's file format, and the DEBUG-style dump of the swept table."""
    exe_dir = BIN / prog
    for name in ("GPU_Hyperlane_Share", "GPU_Hyperlane_Cache", "generator"):
        assert (exe_dir / name).exists(), f"{exe_dir / name} not built"
    # The reference generator stops at 2^13; write a small file by hand in the same format.
    n1, n2 = 64, 32
    t = _table(n1, n2 + 1, 21)  # n2 + 3 rows
    (tmp_path / "Data").mkdir()
    (tmp_path / "Output").mkdir()
    with open(tmp_path / "Data" / "x_2_6_y_2_5.txt", "w") as f:
        f.write(f"{n1} {n2}\n")
        for row in t:
            f.write(" ".join(str(v) for v in row) + " \n")
    print(f"  drivers {prog}", flush=True)
    out = subprocess.run([str(exe_dir / "GPU_Hyperlane_Share"), "6", "5", "32", "128", "3"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert re.fullmatch(r"execution time: [0-9.e+-]+ ms\.\n", out.stdout), out.stdout
    out = subprocess.run([str(exe_dir / "GPU_Hyperlane_Cache"), "6", "5"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert re.fullmatch(r"execution time: [0-9.e+-]+ ms\.\n", out.stdout), out.stdout
    env = dict(os.environ, WFDP_DEBUG_TABLE="1")
    out = subprocess.run([str(exe_dir / "GPU_Hyperlane_Share"), "6", "5", "32", "128", "1"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = (tmp_path / "Output" / f"output_{prog}_6.txt").read_text().splitlines()
    assert int(lines[0]) == n1
    got = np.array([[int(v) for v in ln.split()] for ln in lines[1:]], dtype=np.int32)
    want = O.stencil(op, t[: n2 + 2])
    _check(got, want[1 : n2 + 1, 1 : n1 + 1], f"{prog} driver dump")
    # Usage errors exit with the reference's message.
    out = subprocess.run([str(exe_dir / "GPU_Hyperlane_Share"), "6"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=60)
    assert out.returncode != 0 and "Incorrect Input Parameters" in out.stdout


@pytest.mark.parametrize("op", ["SAT", "SOR"])
def test_stencil_generator_size_matches_oracle(ctx, op):
    """The smallest table the reference's generator writes (2^13 x 2^13,
    SAT/generator.cpp:45), every cell against the oracle."""
    n = 8192
    t = _table(n, n, 13)
    print(f"  stencil {op} {n}x{n}", flush=True)
    r, got = ctx.stencil(op, t)
    _check(got, O.stencil(op, t), f"{op} {n}x{n}")
    assert r.last == got[n, n]
