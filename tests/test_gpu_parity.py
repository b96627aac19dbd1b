"""GPU parity: the HIP engine (through the C ABI) against the oracle on the
same seeded inputs.  Integer work, so the bar is bit-exact."""
import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd import Job, Tile, tile_for, variants

pytestmark = pytest.mark.gpu

VARIANTS = variants()


def rand_pair(rng, n1, n2, alpha):
    return rng.integers(0, alpha, n1, dtype=np.int32), rng.integers(0, alpha, n2, dtype=np.int32)


def check(ctx, prob, a, b, tile=None, textbook=False):
    r = ctx.align(prob, a, b, tile=tile, ed_textbook=textbook)
    last, mx = O.score(prob, a, b, textbook=textbook)
    print(f"  {prob} n1={a.size} n2={b.size} tile={tile} -> {r.last}/{r.max_score} oracle {last}/{mx}", flush=True)
    assert r.status == 0
    assert r.last == last
    if prob == "SW":
        assert r.max_score == mx


@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
def test_small_square_default(ctx, prob):
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 7, 63, 64, 65, 100, 257, 1000):
        a, b = rand_pair(rng, n, n, 4)
        check(ctx, prob, a, b)


@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
def test_non_square(ctx, prob):
    rng = np.random.default_rng(12)
    for n1, n2 in [(1, 300), (300, 1), (5, 700), (700, 5), (129, 1000), (1000, 129), (3000, 777), (64, 4096)]:
        a, b = rand_pair(rng, n1, n2, 4)
        check(ctx, prob, a, b)


@pytest.mark.parametrize("v", VARIANTS, ids=lambda v: f"{v['engine']}-R{v['rows_per_lane']}W{v['waves']}K{v['chunk']}")
@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
def test_every_variant(ctx, prob, v):
    if v["engine"].startswith("bitpar") and prob == "SW":
        pytest.skip("the bit-parallel engine takes LCS and ED only (SW falls back to a cell engine)")
    R, W, K = v["rows_per_lane"], v["waves"], v["chunk"]
    rng = np.random.default_rng(100 + R * 31 + W * 7 + K)
    t = tile_for(v)
    # Several strips per pair (strip-to-strip exchange through HBM), ragged
    # edges, a pair narrower than one wave's lag, and one longer than the
    # LDS rings and code window.
    for n1, n2 in [(777, 64 * W * R * 3 + 17), (100, 64 * W * R + 1), (1500, 1333), (5000, 64 * W * R * 2),
                   (10, 64 * W * R * 2 + 5)]:
        a, b = rand_pair(rng, n1, n2, 4)
        check(ctx, prob, a, b, tile=t)
    plan = ctx.plan([Job(prob, a, b)], t)
    assert plan.describe()["engine"] == v["engine"] and plan.describe()["waves"] == W
    plan.close()


@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
@pytest.mark.parametrize("alpha", [2, 8, 9, 20, 32, 33, 200])
def test_alphabets(ctx, prob, alpha):
    rng = np.random.default_rng(alpha)
    a, b = rand_pair(rng, 900, 1100, alpha)
    check(ctx, prob, a, b)


@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
def test_arbitrary_symbols(ctx, prob):
    rng = np.random.default_rng(5)
    syms = np.array([-7, 2**31 - 1, -2**31, 12345, 0, 99], dtype=np.int32)
    a = syms[rng.integers(0, syms.size, 800)]
    b = syms[rng.integers(0, syms.size, 600)]
    check(ctx, prob, a, b)


BITPAR = Tile(engine="bitpar")
BITPAR_STAGED = Tile(engine="bitpar_staged")


@pytest.mark.parametrize("tile", [BITPAR, BITPAR_STAGED], ids=["plain", "staged"])
@pytest.mark.parametrize("prob,textbook", [("LCS", False), ("ED", False), ("ED", True)])
def test_bitpar_word_and_wave_edges(ctx, prob, textbook, tile):
    """The bit-parallel engine at the edges of its geometry: rows around one
    32-row word, one 64-lane wave (2048 rows) and a 4-wave strip, columns
    around one wave's 64-step skew, 2- to 32-letter alphabets."""
    rng = np.random.default_rng(31)
    for n1, n2, alpha in [(1, 1, 4), (2, 31, 2), (3, 32, 4), (64, 33, 4), (63, 2047, 20), (65, 2048, 32),
                          (200, 2049, 4), (1, 8193, 4), (2500, 8192 * 2 + 3, 20), (4097, 100, 3)]:
        a, b = rand_pair(rng, n1, n2, alpha)
        check(ctx, prob, a, b, tile=tile, textbook=textbook)


def test_bitpar_is_the_default_for_lcs_and_ed(ctx):
    rng = np.random.default_rng(32)
    a, b = rand_pair(rng, 3000, 5000, 20)
    for prob in ("LCS", "ED"):
        # One pair of two strips: the chip is underfilled, so the staged variant.
        plan = ctx.plan([Job(prob, a, b)])
        assert plan.describe()["engine"] == "bitpar_staged"
        plan.close()
        # Hundreds of pairs: more strips than CUs, the plain variant.
        plan = ctx.plan([Job(prob, a[:500], b[:300])] * 600)
        assert plan.describe()["engine"] == "bitpar"
        plan.close()
    for prob, alpha in (("SW", 4), ("ED", 40)):  # SW, and more than 32 symbols: cell engines
        a2, b2 = rand_pair(rng, 500, 700, alpha)
        plan = ctx.plan([Job(prob, a2, b2)])
        assert plan.describe()["engine"] != "bitpar"
        plan.close()
    from gpumemefficient_amd import WfError
    with pytest.raises(WfError):
        ctx.align("SW", a, b, tile=BITPAR)


def test_ed_textbook(ctx):
    rng = np.random.default_rng(6)
    for n1, n2 in [(1, 1), (2, 2), (50, 70), (1000, 999)]:
        a, b = rand_pair(rng, n1, n2, 4)
        check(ctx, "ED", a, b, textbook=True)
    # Known answer: arr1=[1,2], arr2=[2,3] -> ED.cpp 1, textbook 2.
    a, b = np.array([1, 2], np.int32), np.array([2, 3], np.int32)
    assert ctx.align("ED", a, b).last == 1
    assert ctx.align("ED", a, b, ed_textbook=True).last == 2


@pytest.mark.parametrize("prob", ["LCS", "ED", "SW"])
def test_empty(ctx, prob):
    e = np.zeros(0, np.int32)
    x = np.arange(5, dtype=np.int32)
    for a, b in [(e, e), (e, x), (x, e)]:
        check(ctx, prob, a, b)


def test_identity_closed_forms(ctx):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 4, 5000, dtype=np.int32)
    assert ctx.align("LCS", a, a).last == 5000
    assert ctx.align("ED", a, a).last == 0
    r = ctx.align("SW", a, a)
    assert r.last == 3 * 5000 and r.max_score == 3 * 5000


def test_batch_mixed(ctx):
    rng = np.random.default_rng(8)
    jobs = []
    for i in range(200):
        prob = ["LCS", "ED", "SW"][i % 3]
        n1, n2 = int(rng.integers(0, 1500)), int(rng.integers(0, 1500))
        a, b = rand_pair(rng, n1, n2, 4 if prob != "ED" else 20)
        jobs.append(Job(prob, a, b))
    res = ctx.align_batch(jobs)
    for j, r in zip(jobs, res):
        last, mx = O.score(j.problem, j.a, j.b)
        assert r.status == 0
        assert (r.last, r.max_score if j.problem == "SW" else 0) == (last, mx)


def test_plan_persistent_equals_per_pair(ctx):
    rng = np.random.default_rng(9)
    jobs = [Job("SW", *rand_pair(rng, int(rng.integers(64, 2000)), int(rng.integers(64, 2000)), 4))
            for _ in range(40)]
    plan = ctx.plan(jobs)
    plan.run()
    r1 = plan.results()
    plan.run(per_pair=True)
    r2 = plan.results()
    for j, x, y in zip(jobs, r1, r2):
        last, mx = O.sw(j.a, j.b)
        assert (x.last, x.max_score) == (last, mx)
        assert (y.last, y.max_score) == (last, mx)
    plan.close()


def _mixed_jobs(seed, count, maxlen):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        prob = ["LCS", "ED", "SW"][i % 3]
        n1, n2 = int(rng.integers(1, maxlen)), int(rng.integers(1, maxlen))
        jobs.append(Job(prob, *rand_pair(rng, n1, n2, 20 if prob == "ED" else 4)))
    return jobs


def _expected(jobs):
    return [O.score(j.problem, j.a, j.b) for j in jobs]


def test_graph_replay_matches_eager(ctx):
    """hipGraph replays of a captured plan, interleaved with eager work of
    another plan and host synchronisation, all equal the eager result."""
    jobs = _mixed_jobs(21, 60, 2500)
    want = _expected(jobs)
    other = ctx.plan(_mixed_jobs(22, 10, 800))
    for per_pair in (False, True):
        plan = ctx.plan(jobs)
        plan.run(per_pair=per_pair)
        eager = [(r.last, r.max_score) for r in plan.results()]
        assert eager == want
        plan.capture(per_pair=per_pair)
        for _ in range(5):
            plan.run(per_pair=per_pair)
            assert [(r.last, r.max_score) for r in plan.results()] == want
            other.run()
            other.results()
        plan.close()
    other.close()


def test_autotune_roundtrip(ctx, tmp_path):
    from gpumemefficient_amd.api import tune_clear, tune_load, tune_save
    jobs = _mixed_jobs(23, 24, 1200)
    tile, ms = ctx.autotune(jobs, reps=1)
    assert ms > 0 and tile.engine in ("lockstep", "async")
    path = tmp_path / "tune.tsv"
    tune_save(path)
    tune_clear()
    tune_load(path)
    plan = ctx.plan(jobs)           # no tile: the tuned choice
    d = plan.describe()
    assert d["engine"] == tile.engine and d["chunk"] == tile.tile_w
    assert d["rows_per_lane"] * d["waves"] * 64 == tile.strip_rows
    plan.run()
    assert [(r.last, r.max_score) for r in plan.results()] == _expected(jobs)
    plan.close()
    tune_clear()


def test_sw_byte_tables_match_field_tables(ctx, monkeypatch):
    """SW pairs of at most 4 symbols on the asynchronous engine with two rows
    per lane use byte row tables and one v_perm_b32 per step; the same
    pairs with the 4-bit field tables (WFDP_PERM=0) and the oracle agree
    on the last and the maximum cell: odd row counts (a padding row in the
    last lane), pairs narrower than a wave's skew, 1 to 4 symbols."""
    rng = np.random.default_rng(91)
    t = Tile(engine="async", rows_per_lane=2, strip_rows=512)
    cases = [(1, 1, 1), (5, 3, 2), (63, 129, 4), (700, 1001, 4), (2000, 515, 3), (130, 4097, 4), (3000, 3000, 1)]
    for n1, n2, alpha in cases:
        a, b = rand_pair(rng, n1, n2, alpha)
        want = O.score("SW", a, b)
        got = {}
        for flag in ("1", "0"):
            monkeypatch.setenv("WFDP_PERM", flag)
            r = ctx.align("SW", a, b, tile=t)
            got[flag] = (r.last, r.max_score)
        print(f"  SW perm n1={n1} n2={n2} alpha={alpha}: {got} oracle {want}", flush=True)
        assert got["1"] == got["0"] == tuple(want)
