"""The headline workloads at full size, bit-exact: against the oracle's
answers stored by tests/golden/make_fullsize.py (the oracle needs about a
minute of CPU for them), and through size-independent properties -- the
LCS and SW scores are symmetric in the two sequences, textbook edit
distance too, both engines agree -- and every result of the full served
trace, in both serving modes, equals the oracle's."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from gpumemefficient_amd import Job, Tile
from gpumemefficient_amd import datagen as G

FIX = json.loads((Path(__file__).parent / "golden" / "fullsize.json").read_text())


def _digest(*arrays):
    h = hashlib.sha256()
    for x in arrays:
        h.update(np.ascontiguousarray(x, dtype=np.int32).tobytes())
    return h.hexdigest()


def test_seeded_inputs_unchanged():
    """The fixtures only mean something for the same inputs (CPU test)."""
    assert _digest(*G.lcs_pair()) == FIX["lcs_8192"]["inputs_sha256"]
    assert _digest(*G.ed_pair()) == FIX["ed_65536"]["inputs_sha256"]
    assert _digest(*G.sw_pair()) == FIX["sw_131072"]["inputs_sha256"]
    assert _digest(*[x for p in G.batch_pairs() for x in p]) == FIX["sw_batch_4096"]["inputs_sha256"]


LOCKSTEP = Tile(tile_w=64, strip_rows=512, rows_per_lane=2, engine="lockstep")


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [None, LOCKSTEP], ids=["default", "lockstep"])
def test_lcs_8192(ctx, tile):
    a, b = G.lcs_pair()
    assert ctx.align("LCS", a, b, tile).last == FIX["lcs_8192"]["last"]
    assert ctx.align("LCS", b, a, tile).last == FIX["lcs_8192"]["last"]  # symmetric


@pytest.mark.gpu
@pytest.mark.parametrize("tx", [128, 256])
def test_lcs_8192_through_the_drop_in_driver(tmp_path, tx):
    """LCS of the 8192 x 8192 seed-1 pair through the drop-in
    LCS/GPU_Hyperlane_Share at tile widths 128 and 256 with 1024 threads per
    tile (the driver's 5-argument form, e1 e2 tX tY itr; the reference's own
    host hard-codes 32 x 256, LCS/GPU_Hyperlane_Share.cu:703-704): the
    printed length is the oracle's."""
    import re
    import subprocess
    from gpumemefficient_amd.datafile import write_input
    exe = Path(__file__).resolve().parent.parent / "gpumemefficient_amd" / "bin" / "LCS" / "GPU_Hyperlane_Share"
    a, b = G.lcs_pair()
    write_input(tmp_path / "Data" / "str1_2_13_str2_2_13.txt", a, b)
    r = subprocess.run([str(exe), "13", "13", str(tx), "1024", "1"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "input data loaded"
    assert lines[1] == f"lcs length: {FIX['lcs_8192']['last']}"
    assert re.fullmatch(r"execution time: [0-9.e+-]+ ms\.", lines[2]), lines[2]


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [None, LOCKSTEP], ids=["default", "lockstep"])
def test_ed_65536(ctx, tile):
    a, b = G.ed_pair()
    assert ctx.align("ED", a, b, tile).last == FIX["ed_65536"]["last"]
    # The textbook distance is symmetric (ED.cpp's row-0 boundary is not).
    t1 = ctx.align("ED", a, b, tile, ed_textbook=True).last
    assert t1 == ctx.align("ED", b, a, tile, ed_textbook=True).last
    # ED.cpp(a, b) = min(lev(a, b), lev(a[1:], b)).
    assert FIX["ed_65536"]["last"] == min(t1, ctx.align("ED", a[1:], b, tile, ed_textbook=True).last)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [None, LOCKSTEP], ids=["default", "lockstep"])
def test_sw_131072(ctx, tile):
    a, b = G.sw_pair()
    r = ctx.align("SW", a, b, tile)
    assert (r.last, r.max_score) == (FIX["sw_131072"]["last"], FIX["sw_131072"]["max"])
    r2 = ctx.align("SW", b, a, tile)
    assert (r2.last, r2.max_score) == (r.last, r.max_score)


@pytest.mark.gpu
def test_sw_batch_4096(ctx):
    jobs = [Job("SW", a, b) for a, b in G.batch_pairs()]
    want = list(zip(FIX["sw_batch_4096"]["last"], FIX["sw_batch_4096"]["max"]))
    plan = ctx.plan(jobs)
    for per_pair in (False, True):
        plan.run(per_pair=per_pair)
        assert [(r.last, r.max_score) for r in plan.results()] == want
    plan.capture()
    for _ in range(3):
        plan.run()
        assert [(r.last, r.max_score) for r in plan.results()] == want
    plan.close()


@pytest.mark.gpu
def test_served_trace_equals_oracle(ctx):
    """The full 600-request served mix (lengths up to 131072, 7.6e10 cells)
    in both serving modes, replayed at 50 000 requests/s: every result
    equals the oracle's (its all-cores variant, anti-diagonals of 256 x 256
    tiles, a few seconds on 16 host threads) and the per-pair launch's."""
    import os
    import oracle as O
    from gpumemefficient_amd.server import Server
    reqs = G.serving_requests(50000.0, 600, seed=5)
    jobs = [Job(r.problem, r.a, r.b) for r in reqs]
    threads = max(1, min(16, os.cpu_count() or 1))
    want = [O.blocked(j.problem, j.a, j.b, threads, 256) for j in jobs]
    for resident in (False, True):
        with Server(0, resident=resident) as srv:
            comps = srv.replay(jobs, [r.arrival_s for r in reqs])
        print(f"  resident={resident}: {len(comps)} completions", flush=True)
        assert all(c.status == 0 for c in comps)
        assert [(c.last, c.max_score) for c in comps] == want
    plan = ctx.plan(jobs)
    plan.run(per_pair=True)
    assert [(r.last, r.max_score) for r in plan.results()] == want
    plan.close()
