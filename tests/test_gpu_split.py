"""Column-split bit-parallel pairs: an LCS or ED pair of at least 1024 columns
in a launch with room for twice its strips runs as two halves at once (the
left columns forward, the right columns reversed) and the score is combined
on the device from the two halves' last columns.  Bit-exact against the
oracle, against the unsplit engine (WFDP_SPLIT=0), in batches mixed with
unsplit pairs, per-pair launches and graph replays."""
import os

import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd import Job, Tile

pytestmark = pytest.mark.gpu

# (n1, n2): the smallest split width, odd widths, rows around one word, one
# 2048-row strip and several, rows not a multiple of 32, tall and wide.
SHAPES = [(1024, 1), (1025, 31), (1024, 32), (1031, 33), (2047, 2048), (2048, 2049), (3000, 5000),
          (4097, 4100), (20000, 3000), (1500, 70000), (65536, 257)]


def _pair(rng, n1, n2, alpha):
    return rng.integers(0, alpha, n1, dtype=np.int32), rng.integers(0, alpha, n2, dtype=np.int32)


def _tasks_unsplit(n2, strip_rows=2048):
    return (n2 + strip_rows - 1) // strip_rows


@pytest.mark.parametrize("prob,textbook", [("LCS", False), ("ED", False), ("ED", True)])
@pytest.mark.parametrize("n1,n2", SHAPES)
def test_split_matches_oracle(ctx, prob, textbook, n1, n2):
    rng = np.random.default_rng(n1 * 31 + n2)
    for alpha in (2, 4, 20):
        a, b = _pair(rng, n1, n2, alpha)
        want = O.score(prob, a, b, textbook=textbook)[0]
        plan = ctx.plan([Job(prob, a, b, ed_textbook=textbook)])
        d = plan.describe()
        plan.run()
        r = plan.results()[0]
        plan.close()
        print(f"  split {prob} tb={textbook} n1={n1} n2={n2} alpha={alpha}: {r.last} oracle {want} "
              f"tasks={d['tasks']}", flush=True)
        assert d["engine"].startswith("bitpar")
        assert d["tasks"] == 2 * _tasks_unsplit(n2), "expected the split path"
        assert r.status == 0 and r.last == want


def test_split_off_gives_the_same_scores(ctx, monkeypatch):
    rng = np.random.default_rng(77)
    jobs = [Job("ED", *_pair(rng, 5000, 6000, 20)), Job("LCS", *_pair(rng, 4096, 3000, 4))]
    got = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("WFDP_SPLIT", flag)
        plan = ctx.plan(jobs)
        d = plan.describe()
        plan.run()
        got[flag] = [r.last for r in plan.results()]
        plan.close()
        nt = sum(_tasks_unsplit(j.b.size) for j in jobs)
        assert d["tasks"] == (2 * nt if flag == "1" else nt)
    assert got["1"] == got["0"]
    assert got["1"] == [O.score(j.problem, j.a, j.b)[0] for j in jobs]


def test_split_in_a_mixed_batch(ctx):
    """Split pairs next to pairs too narrow to split, empty pairs and both
    problems, in one persistent launch and in per-pair launches."""
    rng = np.random.default_rng(78)
    jobs = []
    for k in range(12):
        prob = ("ED", "LCS")[k % 2]
        n1 = int(rng.integers(1, 3000)) if k % 3 else 0
        n2 = int(rng.integers(1, 4000))
        jobs.append(Job(prob, *_pair(rng, n1, n2, 4 + k)))
    want = [O.score(j.problem, j.a, j.b)[0] for j in jobs]
    plan = ctx.plan(jobs)
    for per_pair in (False, True):
        plan.run(per_pair=per_pair)
        got = [r.last for r in plan.results()]
        print(f"  mixed batch per_pair={per_pair}: {got}", flush=True)
        assert got == want
    plan.close()


def test_split_graph_replays(ctx):
    """The engine launch and the combine kernel captured into one hipGraph:
    every replay, with eager work between replays, gives the eager scores."""
    rng = np.random.default_rng(79)
    jobs = [Job("ED", *_pair(rng, 7000, 9000, 20)), Job("LCS", *_pair(rng, 3000, 2500, 4))]
    want = [O.score(j.problem, j.a, j.b)[0] for j in jobs]
    plan = ctx.plan(jobs)
    plan.run()
    assert [r.last for r in plan.results()] == want
    plan.capture()
    for rep in range(4):
        plan.run()
        got = [r.last for r in plan.results()]
        print(f"  replay {rep}: {got}", flush=True)
        assert got == want
        a, b = _pair(rng, 1500, 1700, 4)
        assert ctx.align("LCS", a, b).last == O.score("LCS", a, b)[0]
    plan.close()


def test_split_is_skipped_when_the_chip_is_full(ctx):
    """Hundreds of pairs: more strips than half the CUs, no split."""
    rng = np.random.default_rng(80)
    a, b = _pair(rng, 2000, 2500, 4)
    plan = ctx.plan([Job("ED", a, b)] * 300)
    d = plan.describe()
    plan.run()
    rs = plan.results()
    plan.close()
    assert d["tasks"] == 300 * _tasks_unsplit(2500, 64 * d["waves"] * 32)
    assert all(r.last == rs[0].last for r in rs)
    assert rs[0].last == O.score("ED", a, b)[0]


def test_forced_engine_tiles_still_split(ctx):
    """Explicit bit-parallel tiles (either variant, any strip height) split
    as well; scores stay exact."""
    rng = np.random.default_rng(81)
    a, b = _pair(rng, 9000, 5000, 20)
    want = O.score("ED", a, b)[0]
    for t in (Tile(engine="bitpar"), Tile(engine="bitpar_staged"), Tile(engine="bitpar", strip_rows=8192),
              Tile(engine="bitpar_staged", strip_rows=4096)):
        r = ctx.align("ED", a, b, tile=t)
        print(f"  {t}: {r.last} oracle {want}", flush=True)
        assert r.status == 0 and r.last == want


@pytest.mark.skipif(os.environ.get("WFDP_SPLIT") == "0", reason="split disabled in the environment")
def test_split_full_size_ed_pair(ctx):
    """The 65536 x 65536 edit-distance pair (20 letters, seed 2), split:
    the distance the reference CPU program gives on this input
    (tests/golden/fullsize.json)."""
    import json
    from pathlib import Path
    from gpumemefficient_amd import datagen as G
    want = json.loads((Path(__file__).parent / "golden" / "fullsize.json").read_text())["ed_65536"]["last"]
    a, b = G.ed_pair()
    plan = ctx.plan([Job("ED", a, b)])
    plan.run()
    r = plan.results()[0]
    d = plan.describe()
    plan.close()
    assert d["tasks"] == 2 * _tasks_unsplit(65536)
    assert r.last == want


TIMEOUT_CASES = [
    # (problem, n1, n2, alphabet, tile): the two bit-parallel engines on a
    # multi-strip ED pair; the asynchronous and lock-step cell engines on SW
    # pairs, one of them a single strip (no output wave, and an input wave
    # that stages all 200 columns at once and leaves).
    ("ED", 20000, 9000, 20, None),
    ("ED", 20000, 9000, 20, Tile(engine="bitpar")),
    ("SW", 200, 300, 4, Tile(engine="async")),
    ("SW", 3000, 2000, 4, Tile(engine="async")),
    ("SW", 200, 300, 4, Tile(engine="lockstep")),
    ("SW", 3000, 2000, 4, Tile(engine="lockstep")),
    ("SW", 3000, 2000, 4, Tile(engine="async", no_max=True)),
]


@pytest.mark.parametrize("prob,n1,n2,alpha,tile", TIMEOUT_CASES,
                         ids=["ed-staged", "ed-plain", "sw-async-1strip", "sw-async", "sw-lockstep-1strip",
                              "sw-lockstep", "sw-async-nomax"])
def test_dependency_timeout_is_reported_and_the_next_run_is_exact(prob, n1, n2, alpha, tile):
    """A context whose dependency waits may last 1 us: waves and strips
    that wait for the ones above give up, the launch still drains (every
    wait is bounded) and either reports a device-side timeout or, if no
    wait ran out, returns the exact score -- never a wrong score with
    status ok.  A context with the normal limit then gives the exact
    score."""
    from gpumemefficient_amd import Context, WfError
    rng = np.random.default_rng(82)
    a, b = _pair(rng, n1, n2, alpha)
    want = O.score(prob, a, b)
    with Context(0, timeout_s=1e-6) as c:
        try:
            r = c.align(prob, a, b, tile=tile)
            print(f"  no wait exceeded 1 us: {r.last}", flush=True)
            assert r.last == want[0]
            if prob == "SW":
                assert r.max_score == (-1 if tile is not None and tile.no_max else want[1])
        except WfError as e:
            print(f"  timeout reported: {e}", flush=True)
            from gpumemefficient_amd import _lib
            assert e.status == _lib.WF_ERR_TIMEOUT
    with Context(0, timeout_s=20.0) as c:
        r = c.align(prob, a, b, tile=tile)
        assert r.status == 0 and r.last == want[0]
        if prob == "SW":
            assert r.max_score == (-1 if tile is not None and tile.no_max else want[1])
