"""SW without the running maximum (WF_TILE_NO_MAX): the bottom-right cell
only, which is all the reference's SW() returns (SmithWaterman/SW.cpp:80).
The cell engines run instances without the maximum; the bottom-right cell
must equal the oracle's, and max_score reads -1, on every engine and table
kind.  Pairs whose alphabet exceeds the row tables keep the maximum.  Also:
an engine named in a tile is honoured even where the autotune table has an
entry for the workload."""
import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd import Job, Tile

pytestmark = pytest.mark.gpu


def pair(rng, n1, n2, alpha):
    return rng.integers(0, alpha, n1, dtype=np.int32), rng.integers(0, alpha, n2, dtype=np.int32)


@pytest.mark.parametrize("engine", ["async", "lockstep"])
@pytest.mark.parametrize("alpha", [4, 6, 20])
def test_no_max_last_equals_oracle(ctx, engine, alpha):
    """alpha 4: byte tables (two rows per lane); 6: 4-bit field tables; 20:
    compared codes (when a pair uses more than 8 symbols), which keep the
    maximum."""
    rng = np.random.default_rng(40 + alpha)
    for n1, n2 in [(1, 1), (63, 700), (1000, 1000), (2049, 1333), (4096, 64)]:
        a, b = pair(rng, n1, n2, alpha)
        last, mx = O.score("SW", a, b)
        for t in (Tile(engine=engine, no_max=True), Tile(engine=engine, rows_per_lane=1, no_max=True)):
            r = ctx.align("SW", a, b, tile=t)
            print(f"  SW a={alpha} {n1}x{n2} {t} -> {r.last}/{r.max_score} oracle {last}/{mx}", flush=True)
            assert r.status == 0 and r.last == last
            # More than 8 distinct symbols: codes are compared, the maximum kept.
            compared = len(set(a.tolist()) | set(b.tolist())) > 8
            assert r.max_score == (mx if compared else -1)


def test_no_max_job_flag_in_a_mixed_batch(ctx):
    """The flag per job: in one launch, no_max pairs report -1 and the others
    their maximum; every bottom-right cell equals the oracle's."""
    rng = np.random.default_rng(47)
    jobs = []
    for i in range(24):
        a, b = pair(rng, int(rng.integers(50, 1500)), int(rng.integers(50, 1500)), 4)
        jobs.append(Job("SW", a, b, no_max=bool(i % 2)))
    for tile in (None, Tile(engine="async"), Tile(engine="lockstep")):
        res = ctx.align_batch(jobs, tile)
        for j, r in zip(jobs, res):
            last, mx = O.score("SW", j.a, j.b)
            assert r.status == 0 and r.last == last
            assert r.max_score == (-1 if j.no_max else mx)


def test_no_max_does_not_touch_lcs_and_ed(ctx):
    rng = np.random.default_rng(48)
    a, b = pair(rng, 900, 1100, 4)
    for prob in ("LCS", "ED"):
        r = ctx.align(prob, a, b, tile=Tile(no_max=True))
        assert r.last == O.score(prob, a, b)[0]


def test_reference_signature_sw_returns_last(ctx):
    """SW(n1, n2, arr1, arr2, paddX, paddY, table, tX, tY) skips the maximum
    and returns the bottom-right cell."""
    from gpumemefficient_amd import SW
    rng = np.random.default_rng(49)
    a, b = pair(rng, 3000, 2500, 4)
    assert SW(a.size, b.size, a, b, 3, 1, None, 32, 128) == O.score("SW", a, b)[0]


def test_engine_flag_beats_the_autotune_table(ctx):
    """Sizes within two octaves of a table entry (ED 2^16 -> bit-parallel,
    SW 2^17 -> async) still run the engine a tile names."""
    rng = np.random.default_rng(50)
    a, b = pair(rng, 16384, 16384, 4)
    p = ctx.plan([Job("ED", a, b)], Tile(engine="async"))
    assert p.describe()["engine"] == "async"
    p.close()
    a, b = pair(rng, 32768, 32768, 4)
    p = ctx.plan([Job("SW", a, b)], Tile(engine="lockstep"))
    assert p.describe()["engine"] == "lockstep"
    p.close()
