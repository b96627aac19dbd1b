"""CPU model of one wave of the wavefront engines (64 lanes, R rows per lane,
lane l at column s - l in step s, the row above lane 0 from the top boundary,
the row above lane l >= 1 from lane l-1 one step earlier), used to check the
free-head start states of the asynchronous engine against the oracle:

- LCS and ED lanes left of column 0 run the ordinary step from a sentinel
  (-2^29 for LCS, +2^29 for ED in the E = D - x - y form) and reach the left
  boundary 0 exactly at column -1;
- SW lanes start at 0 and read a column code that matches no row.

A wrong sentinel or a matching code left of column 0 must fail the check."""
import random

import numpy as np
import pytest

import oracle as O

BIG = 1 << 29
IDLE = {"LCS": -BIG, "ED": BIG, "SW": 0}


def _cell(prob, up, left, diag, match):
    if prob == "LCS":
        return max(up, left, diag + (1 if match else 0))
    if prob == "ED":
        return min(up, left, diag + (0 if match else 1) - 2)
    return max(max(up, left, diag + (5 if match else -1)) - 2, 0)


def run_wave(prob, a, b, R=2, idle=None, never_code=True, rng=None):
    """Sweep one wave over the pair (n2 <= 64 R) with unmasked steps left of
    column 0 and frozen lanes right of column n1 - 1; returns (last, max)."""
    idle = IDLE[prob] if idle is None else idle
    rng = rng or random.Random(0)
    n1, n2 = len(a), len(b)
    L = 64
    assert n2 <= L * R
    h = [[0] * R] + [[idle] * R for _ in range(L - 1)]
    dg = [0] + [idle] * (L - 1)
    best = 0
    top = -1 if prob == "ED" else 0   # ED.cpp row 0: D(0, x) = x - 1
    for s in range(n1 + L - 1):
        prev = [h[l][R - 1] for l in range(L)]
        for l in range(L):
            x = s - l
            if x < -L or x >= n1:
                continue
            if x >= 0:
                col = a[x]
            elif prob == "SW" and never_code:
                col = None
            else:
                col = rng.randint(0, 3)   # whatever the code window holds
            up = top if l == 0 else prev[l - 1]
            up0, diag = up, dg[l]
            for r in range(R):
                y = l * R + r
                match = col is not None and y < n2 and col == b[y]
                left = h[l][r]
                v = _cell(prob, up, left, diag, match)
                diag, h[l][r], up = left, v, v
                if prob == "SW" and y < n2 and x >= 0:
                    best = max(best, v)
            dg[l] = up0
    y = n2 - 1
    v = h[y // R][y % R]
    if prob == "ED":
        v += n1 + n2
    return v, (best if prob == "SW" else 0)


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    for i in range(count):
        prob = ("LCS", "ED", "SW")[i % 3]
        n1, n2 = int(rng.integers(1, 120)), int(rng.integers(1, 129))
        yield prob, rng.integers(0, 4, n1, dtype=np.int32), rng.integers(0, 4, n2, dtype=np.int32)


def test_free_head_matches_oracle():
    for prob, a, b in _cases(1, 60):
        got = run_wave(prob, list(a), list(b))
        last, mx = O.score(prob, a, b)
        assert got == (last, mx if prob == "SW" else 0), (prob, a.size, b.size)


@pytest.mark.parametrize("prob", ["LCS", "ED"])
def test_wrong_sentinel_is_detected(prob):
    bad = 0
    for p, a, b in _cases(2, 45):
        if p != prob:
            continue
        got = run_wave(p, list(a), list(b), idle=0)
        bad += got != (O.score(p, a, b)[0], 0)
    assert bad > 0


def test_matching_code_left_of_column_0_is_detected():
    bad = 0
    for p, a, b in _cases(3, 45):
        if p != "SW":
            continue
        got = run_wave(p, list(a), list(b), never_code=False, rng=random.Random(5))
        bad += got != O.score(p, a, b)
    assert bad > 0
