"""Column-split SW pairs: an SW pair of at least 1024 columns on the byte
tables (at most 4 symbols), in a launch with room for all of its pairs'
strips, runs as three pairs at once -- the left columns forward, the right
columns forward from a zero boundary, the right columns reversed without
the clamp -- when its output is the bottom-right cell only (no_max, the
reference's SW()), and as four with the running maximum (plus the right
columns reversed with the clamp); the bottom-right cell and the maximum are
combined on the device from the halves' last columns.  Bit-exact against
the oracle and the unsplit engine (WFDP_SPLIT=0), with alignments that
cross the cut, per-pair launches and graph replays."""
import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd import Job, Tile

pytestmark = pytest.mark.gpu

# (n1, n2): the smallest split width, odd widths, one row, rows around a
# wave (128) and a strip (512), several strips, wide and short, tall.
SHAPES = [(1024, 1), (1025, 63), (1500, 128), (2047, 513), (3000, 600), (4096, 1), (4097, 63), (4100, 128),
          (4099, 513), (5000, 1000), (8191, 2047), (9000, 4097), (32768, 600), (4096, 20000), (1024, 20000)]


def _pair(rng, n1, n2, alpha):
    return rng.integers(0, alpha, n1, dtype=np.int32), rng.integers(0, alpha, n2, dtype=np.int32)


def _strips(n2, strip_rows=512):
    return (n2 + strip_rows - 1) // strip_rows


def _expected_tasks(d, n2, pairs=3):
    """Strips of the plan described by d: `pairs` times the pair's when the
    chosen kernel has two rows per lane (the byte tables the split needs;
    shorter pairs get one row per lane), once otherwise."""
    strips = _strips(n2, 64 * d["waves"] * d["rows_per_lane"])
    return pairs * strips if d["rows_per_lane"] == 2 else strips


@pytest.mark.parametrize("n1,n2", SHAPES)
def test_sw_split_matches_oracle(ctx, n1, n2):
    rng = np.random.default_rng(n1 * 17 + n2)
    for alpha in (2, 3, 4):
        a, b = _pair(rng, n1, n2, alpha)
        want = O.score("SW", a, b)[0]
        plan = ctx.plan([Job("SW", a, b, no_max=True)])
        d = plan.describe()
        plan.run()
        r = plan.results()[0]
        plan.close()
        print(f"  SW {n1}x{n2} alpha {alpha}: {d['engine']} tasks {d['tasks']} last {r.last} oracle {want}",
              flush=True)
        assert d["engine"] == "async" and d["tasks"] == _expected_tasks(d, n2)
        assert r.status == 0 and r.last == want and r.max_score == -1
        if n2 >= 512:
            assert d["rows_per_lane"] == 2 and d["tasks"] == 3 * _strips(n2, 128 * d["waves"])


@pytest.mark.parametrize("n1,n2", SHAPES)
def test_sw_split_with_the_maximum_matches_oracle(ctx, n1, n2):
    rng = np.random.default_rng(n1 * 19 + n2)
    for alpha in (2, 4):
        a, b = _pair(rng, n1, n2, alpha)
        want = O.score("SW", a, b)
        plan = ctx.plan([Job("SW", a, b)])
        d = plan.describe()
        plan.run()
        r = plan.results()[0]
        plan.close()
        print(f"  SW {n1}x{n2} alpha {alpha} with max: tasks {d['tasks']} got {(r.last, r.max_score)} "
              f"oracle {want}", flush=True)
        assert d["engine"] == "async" and d["tasks"] == _expected_tasks(d, n2, 4)
        assert r.status == 0 and (r.last, r.max_score) == want


def test_sw_split_maximum_across_the_cut(ctx):
    """The best local alignment crosses the middle column, lies left of it,
    or lies right of it; a repeat on both sides competes."""
    rng = np.random.default_rng(77)
    for n1, n2, lo, hi in [(6000, 3000, 2000, 4000), (6000, 3000, 200, 2500), (6000, 3000, 3500, 5800),
                           (8192, 1024, 4096 - 512, 4096 + 512)]:
        a, b = _pair(rng, n1, n2, 4)
        seg = a[lo:hi].copy()
        seg[::89] = (seg[::89] + 1) % 4
        k = min(len(seg), n2 - 100)
        b[50:50 + k] = seg[:k]
        want = O.score("SW", a, b)
        plan = ctx.plan([Job("SW", a, b)])
        assert plan.describe()["tasks"] == 4 * _strips(n2)
        plan.run()
        r = plan.results()[0]
        plan.close()
        print(f"  planted max {n1}x{n2} [{lo},{hi}): got {(r.last, r.max_score)} oracle {want}", flush=True)
        assert r.status == 0 and (r.last, r.max_score) == want and want[1] > 1000


def test_sw_split_alignment_across_the_cut(ctx):
    """A planted segment makes the best alignment into the corner cross the
    middle column; a second planted segment right of the cut competes."""
    rng = np.random.default_rng(71)
    for n1, n2, lo, hi in [(5000, 4000, 1000, 5000), (6000, 3000, 3500, 6000), (4096, 4096, 0, 4096)]:
        a, b = _pair(rng, n1, n2, 4)
        seg = a[lo:hi].copy()
        seg[::97] = (seg[::97] + 1) % 4  # a few mismatches
        k = min(len(seg), n2)
        b[n2 - k:] = seg[len(seg) - k:]
        want = O.score("SW", a, b)[0]
        plan = ctx.plan([Job("SW", a, b, no_max=True)])
        assert plan.describe()["tasks"] == 3 * _strips(n2)
        plan.run()
        r = plan.results()[0]
        plan.close()
        print(f"  planted {n1}x{n2} [{lo},{hi}): last {r.last} oracle {want}", flush=True)
        assert r.status == 0 and r.last == want and want > 1000


def test_sw_split_off_and_with_the_maximum(ctx, monkeypatch):
    rng = np.random.default_rng(72)
    a, b = _pair(rng, 7000, 3000, 4)
    last, mx = O.score("SW", a, b)
    p = ctx.plan([Job("SW", a, b, no_max=True)])
    p.run()
    assert p.describe()["tasks"] == 3 * _strips(3000) and p.results()[0].last == last
    p.close()
    # The running maximum adds the clamped reverse pair.
    p = ctx.plan([Job("SW", a, b)])
    p.run()
    r = p.results()[0]
    assert p.describe()["tasks"] == 4 * _strips(3000) and (r.last, r.max_score) == (last, mx)
    p.close()
    monkeypatch.setenv("WFDP_SPLIT", "0")
    p = ctx.plan([Job("SW", a, b, no_max=True)])
    p.run()
    assert p.describe()["tasks"] == _strips(3000) and p.results()[0].last == last
    p.close()
    p = ctx.plan([Job("SW", a, b)])
    p.run()
    r = p.results()[0]
    assert p.describe()["tasks"] == _strips(3000) and (r.last, r.max_score) == (last, mx)
    p.close()


def test_sw_split_four_pairs_need_a_cu_per_strip(ctx):
    """86 strips: four pairs (344 strips) exceed the 256 CUs, so the pair
    with the running maximum runs unsplit; bit-exact either way."""
    rng = np.random.default_rng(78)
    a, b = _pair(rng, 4096, 86 * 512, 4)
    last, mx = O.score("SW", a, b)
    p = ctx.plan([Job("SW", a, b)])
    p.run()
    r = p.results()[0]
    assert p.describe()["tasks"] == 86 and (r.last, r.max_score) == (last, mx)
    p.close()
    p = ctx.plan([Job("SW", a, b, no_max=True)])
    p.run()
    assert p.describe()["tasks"] in (86, 3 * 86) and p.results()[0].last == last
    p.close()


def test_sw_split_batch_beyond_a_strip_per_cu_is_not_split(ctx):
    """16 pairs of 4096 x 4096 (128 strips): three pairs each would put two
    strips on a CU, which only a launch of one pair takes; 2 pairs split."""
    rng = np.random.default_rng(80)
    for npairs, split in ((16, False), (2, True)):
        jobs = [Job("SW", *_pair(rng, 4096, 4096, 4), no_max=True) for _ in range(npairs)]
        p = ctx.plan(jobs)
        d = p.describe()
        p.run()
        got = [r.last for r in p.results()]
        p.close()
        assert d["engine"] == "async" and d["tasks"] == (3 if split else 1) * 8 * npairs
        assert got == [O.score("SW", j.a, j.b)[0] for j in jobs]


def test_sw_split_narrower_than_1024_columns_is_not_split(ctx):
    rng = np.random.default_rng(79)
    for nm in (True, False):
        a, b = _pair(rng, 1023, 3000, 4)
        last, mx = O.score("SW", a, b)
        p = ctx.plan([Job("SW", a, b, no_max=nm)])
        p.run()
        r = p.results()[0]
        d = p.describe()
        assert d["tasks"] == _strips(3000, 64 * d["waves"] * d["rows_per_lane"])
        assert (r.last, r.max_score) == (last, -1 if nm else mx)
        p.close()


def test_sw_split_more_than_four_symbols_is_not_split(ctx):
    rng = np.random.default_rng(73)
    a, b = _pair(rng, 6000, 900, 6)
    p = ctx.plan([Job("SW", a, b, no_max=True)])
    p.run()
    assert p.describe()["tasks"] == _strips(900) and p.results()[0].last == O.score("SW", a, b)[0]
    p.close()


def test_sw_split_per_pair_launches_and_graph_replays(ctx):
    rng = np.random.default_rng(74)
    jobs = [Job("SW", *_pair(rng, 4500 + 300 * i, 700 + 250 * i, 4), no_max=i != 1) for i in range(3)]
    want = [O.score("SW", j.a, j.b) for j in jobs]
    want = [(w[0], -1 if i != 1 else w[1]) for i, w in enumerate(want)]
    p = ctx.plan(jobs)
    assert p.describe()["tasks"] == sum((4 if i == 1 else 3) * _strips(j.b.size) for i, j in enumerate(jobs))

    def got():
        return [(r.last, r.max_score) for r in p.results()]

    p.run()
    assert got() == want
    p.run(per_pair=True)
    assert got() == want
    p.capture()
    for rep in range(3):
        p.run()
        assert got() == want
        # Eager work between replays.
        a, b = _pair(rng, 4200, 300 + rep, 4)
        assert ctx.align("SW", a, b, tile=Tile(no_max=True)).last == O.score("SW", a, b)[0]
    p.close()


def test_sw_split_through_the_reference_signature(ctx):
    """SW(n1, n2, arr1, arr2, paddX, paddY, table, tX, tY): the drop-in call
    (bottom-right cell, tile arguments as a hint) takes the split."""
    from gpumemefficient_amd import SW
    rng = np.random.default_rng(75)
    a, b = _pair(rng, 8192, 8192, 4)
    assert SW(a.size, b.size, a, b, 3, 1, None, 32, 128) == O.score("SW", a, b)[0]


def test_sw_split_tile_engine_flag(ctx):
    """An async engine named in a tile still splits; the lock-step engine does not."""
    rng = np.random.default_rng(76)
    a, b = _pair(rng, 5000, 1500, 4)
    want = O.score("SW", a, b)[0]
    for tile, tasks in ((Tile(engine="async", no_max=True), 3 * _strips(1500)),
                        (Tile(engine="async"), 4 * _strips(1500)),
                        (Tile(engine="lockstep", no_max=True), None)):
        r = ctx.align("SW", a, b, tile=tile)
        assert r.status == 0 and r.last == want
        p = ctx.plan([Job("SW", a, b)], tile)
        if tasks is not None:
            assert p.describe()["tasks"] == tasks
        p.close()


def test_sw_split_random_pairs(ctx):
    """40 seeded random pairs of 1024..12000 columns and 1..6000 rows on 2..4
    letters, with and without the maximum, every fourth with a planted
    alignment; each a launch of its own, against the oracle."""
    rng = np.random.default_rng(2025)
    nsplit = 0
    for t in range(40):
        n1 = int(rng.integers(1024, 12001))
        n2 = int(rng.integers(1, 6001))
        while n1 * n2 > 20_000_000:
            n2 //= 2
        alpha = int(rng.integers(2, 5))
        a, b = _pair(rng, n1, max(n2, 1), alpha)
        if t % 4 == 0:
            k = min(b.size, n1 // 3)
            lo, off = int(rng.integers(0, n1 - k + 1)), int(rng.integers(0, b.size - k + 1))
            b[off:off + k] = a[lo:lo + k]
        nm = bool(t % 2)
        last, mx = O.score("SW", a, b)
        p = ctx.plan([Job("SW", a, b, no_max=nm)])
        d = p.describe()
        p.run()
        r = p.results()[0]
        p.close()
        nsplit += d["tasks"] > _strips(b.size, 64 * d["waves"] * d["rows_per_lane"])
        assert r.status == 0 and (r.last, r.max_score) == (last, -1 if nm else mx), (t, n1, b.size, alpha, nm, d)
    assert nsplit >= 30
