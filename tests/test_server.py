"""Serving path: continuous batching, admission by HBM budget, results
identical alone and in a mix (bit-exact against the oracle)."""
import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd import Job
from gpumemefficient_amd import _lib
from gpumemefficient_amd.datagen import serving_requests
from gpumemefficient_amd.server import Server, percentile


def _jobs(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        prob = "ED" if i % 2 else "SW"
        alpha = 20 if prob == "ED" else 4
        n1, n2 = int(rng.integers(lo, hi)), int(rng.integers(lo, hi))
        out.append(Job(prob, rng.integers(0, alpha, n1, dtype=np.int32), rng.integers(0, alpha, n2, dtype=np.int32)))
    return out


def _want(j):
    return O.score(j.problem, j.a, j.b)


def test_percentile_helper():
    assert percentile([1, 2, 3, 4], 50) == 2.5
    assert np.isnan(percentile([], 99))


def test_serving_trace_generator():
    reqs = serving_requests(1000.0, 500, seed=5)
    assert len(reqs) == 500
    t = np.array([r.arrival_s for r in reqs])
    assert np.all(np.diff(t) > 0)
    rate = len(reqs) / t[-1]
    assert 800 < rate < 1250  # Poisson at 1k/s
    probs = {r.problem for r in reqs}
    assert probs == {"ED", "SW"}
    for r in reqs:
        alpha = 20 if r.problem == "ED" else 4
        assert r.a.max() < alpha and r.b.max() < alpha
        assert 100 <= r.a.size <= 131072 and 100 <= r.b.size <= 131072
    # Same seed, same trace.
    again = serving_requests(1000.0, 500, seed=5)
    assert all(np.array_equal(x.a, y.a) and x.arrival_s == y.arrival_s for x, y in zip(reqs, again))


MODES = [pytest.param(False, id="slots"), pytest.param(True, id="resident")]


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_results_and_alone_vs_mix(resident):
    """Every result equals the oracle's, submitted alone (one at a time) and
    in a mix of 120 overlapping requests; with the resident engine the mix
    runs inside one persistent launch that took each request in as it came."""
    jobs = _jobs(31, 120, 50, 3000)
    want = [_want(j) for j in jobs]
    with Server(0, streams=4, resident=resident) as srv:
        # Alone: one request at a time.
        alone = []
        for j in jobs[:20]:
            srv.submit(j)
            srv.wait_idle(60)
            alone += srv.poll()
        # In the mix: everything at once, overlapping on the device.
        tickets = [srv.submit(j) for j in jobs]
        srv.wait_idle(120)
        mix = {c.ticket: c for c in srv.poll()}
        st = srv.stats()
    assert len(alone) == 20 and all(c.status == 0 for c in alone)
    for c, w in zip(alone, want[:20]):
        assert (c.last, c.max_score) == w
    for t, w, j in zip(tickets, want, jobs):
        c = mix[t]
        assert c.status == 0
        assert (c.last, c.max_score) == w, (j.problem, j.a.size, j.b.size)
        assert c.t_submit <= c.t_start <= c.t_done
    assert st["completed"] == 140 and st["hbm_used"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_admission_by_budget(resident):
    """A tight HBM budget forces requests to wait for earlier ones to
    finish; all still complete correctly.  A request that can never fit
    fails with WF_ERR_NOMEM instead of blocking the queue."""
    jobs = _jobs(32, 30, 1000, 4000)
    want = [_want(j) for j in jobs]
    with Server(0, streams=2, hbm_budget=600_000, resident=resident) as srv:
        tickets = [srv.submit(j) for j in jobs]
        big = Job("SW", np.zeros(40000, np.int32), np.zeros(40000, np.int32))
        tb = srv.submit(big)
        srv.wait_idle(120)
        got = {c.ticket: c for c in srv.poll()}
        st = srv.stats()
    for t, w in zip(tickets, want):
        assert got[t].status == 0 and (got[t].last, got[t].max_score) == w
    assert got[tb].status == _lib.WF_ERR_NOMEM
    if not resident:
        assert st["launches"] >= 3  # could not all go in one launch
    assert st["hbm_used"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_replay_poisson(resident):
    reqs = serving_requests(2000.0, 150, seed=7)
    rng = np.random.default_rng(1)
    jobs = []
    for r in reqs:  # cap lengths so the oracle check stays fast
        a, b = r.a[:2500], r.b[:2500]
        jobs.append(Job(r.problem, a, b))
    arrivals = [r.arrival_s for r in reqs]
    with Server(0, resident=resident) as srv:
        comps = srv.replay(jobs, arrivals)
    for c, j, t in zip(comps, jobs, arrivals):
        assert c.status == 0
        assert (c.last, c.max_score) == _want(j)
        assert c.t_submit == t and c.t_done >= c.t_submit
    lat = [c.latency for c in comps]
    assert percentile(lat, 99) < 5.0


@pytest.mark.gpu
def test_server_per_request_baseline_matches():
    jobs = _jobs(33, 20, 50, 2000)
    with Server(0, per_request=True) as srv:
        comps = srv.replay(jobs, [0.0] * len(jobs))
        st = srv.stats()
    assert st["launches"] == len(jobs)
    for c, j in zip(comps, jobs):
        assert (c.last, c.max_score) == _want(j)


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_refuses_a_bad_request_and_serves_the_rest(resident):
    """A request of 300 distinct symbols (more than the one-byte codes hold)
    is refused at submission with WF_ERR_ALPHABET; the good requests around
    it complete exactly, and the server keeps serving afterwards."""
    from gpumemefficient_amd import WfError
    jobs = _jobs(34, 40, 50, 2000)
    want = [_want(j) for j in jobs]
    bad = Job("SW", np.arange(300, dtype=np.int32), np.arange(300, dtype=np.int32)[::-1].copy())
    with Server(0, streams=2, resident=resident) as srv:
        tickets = [srv.submit(j) for j in jobs[:20]]
        with pytest.raises(WfError) as ei:
            srv.submit(bad)
        assert ei.value.status == _lib.WF_ERR_ALPHABET
        tickets += [srv.submit(j) for j in jobs[20:]]
        srv.wait_idle(120)
        got = {c.ticket: c for c in srv.poll()}
        again = srv.submit(jobs[0])
        srv.wait_idle(60)
        more = srv.poll()
    for t, w in zip(tickets, want):
        assert got[t].status == 0 and (got[t].last, got[t].max_score) == w
    assert len(more) == 1 and more[0].ticket == again and (more[0].last, more[0].max_score) == want[0]


@pytest.mark.gpu
def test_resident_server_empty_and_textbook_and_large_alphabets():
    """Requests with an empty side (answered from the boundary), ED with the
    textbook row 0, and pairs of 40 and 256 distinct symbols (compared
    codes) through the resident engine, against the oracle; then a second
    server in the same process (a fresh launch) gives the same answers."""
    rng = np.random.default_rng(35)
    jobs = [Job("ED", np.arange(7, dtype=np.int32), np.zeros(0, np.int32)),
            Job("ED", np.zeros(0, np.int32), np.arange(5, dtype=np.int32)),
            Job("SW", np.zeros(0, np.int32), np.zeros(0, np.int32)),
            Job("ED", rng.integers(0, 20, 900, dtype=np.int32), rng.integers(0, 20, 1100, dtype=np.int32),
                ed_textbook=True),
            Job("ED", rng.integers(0, 40, 700, dtype=np.int32), rng.integers(0, 40, 600, dtype=np.int32)),
            Job("SW", rng.integers(0, 256, 800, dtype=np.int32), rng.integers(0, 256, 900, dtype=np.int32)),
            Job("LCS", rng.integers(-5, 5, 1300, dtype=np.int32), rng.integers(-5, 5, 513, dtype=np.int32))]
    want = [O.score(j.problem, j.a, j.b, textbook=j.ed_textbook) for j in jobs]
    for _ in range(2):
        with Server(0, resident=True) as srv:
            tickets = [srv.submit(j) for j in jobs]
            srv.wait_idle(60)
            got = {c.ticket: c for c in srv.poll()}
        for t, w, j in zip(tickets, want, jobs):
            assert got[t].status == 0
            assert (got[t].last, got[t].max_score) == (w[0], w[1] if j.problem == "SW" else 0), j.problem


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_no_max_requests(resident):
    """SW requests with the no_max job flag: the bottom-right cell equals the
    oracle's and the maximum is skipped (max_score -1) in both modes."""
    rng = np.random.default_rng(34)
    jobs = [Job("SW", rng.integers(0, 4, int(rng.integers(100, 2500)), dtype=np.int32),
                rng.integers(0, 4, int(rng.integers(100, 2500)), dtype=np.int32), no_max=True) for _ in range(16)]
    with Server(0, streams=2, resident=resident) as srv:
        tickets = [srv.submit(j) for j in jobs]
        srv.wait_idle(120)
        got = {c.ticket: c for c in srv.poll()}
    for t, j in zip(tickets, jobs):
        last, mx = O.score("SW", j.a, j.b)
        assert got[t].status == 0 and got[t].last == last
        assert got[t].max_score == -1


@pytest.mark.gpu
@pytest.mark.parametrize("resident", MODES)
def test_server_sw_requests_that_column_split(resident):
    """SW requests of at least 1024 columns on 4 letters, which a launch of
    their own plans as column-split pairs: alone and together, with and
    without the maximum, among them one (20000 rows with the maximum) whose
    four split pairs would exceed a slot's share of the chip.  Bit-exact."""
    rng = np.random.default_rng(35)
    shapes = [(4096, 1000, True), (5000, 3000, False), (4096, 20000, False), (8192, 700, False), (6000, 6000, True),
              (1024, 900, False), (1500, 1500, True)]
    jobs = [Job("SW", rng.integers(0, 4, n1, dtype=np.int32), rng.integers(0, 4, n2, dtype=np.int32), no_max=nm)
            for n1, n2, nm in shapes]
    want = [O.score("SW", j.a, j.b) for j in jobs]
    want = [(w[0], -1 if j.no_max else w[1]) for w, j in zip(want, jobs)]
    with Server(0, streams=4, resident=resident) as srv:
        alone = []
        for j in jobs:
            srv.submit(j)
            srv.wait_idle(60)
            alone += srv.poll()
        tickets = [srv.submit(j) for j in jobs]
        srv.wait_idle(120)
        mix = {c.ticket: c for c in srv.poll()}
    assert [(c.status, c.last, c.max_score) for c in alone] == [(0,) + w for w in want]
    assert [(mix[t].status, mix[t].last, mix[t].max_score) for t in tickets] == [(0,) + w for w in want]
