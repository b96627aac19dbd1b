"""Pins the oracle (oracle/wf_oracle.c) before anything is checked against it.

Each function is tied to something other than its own formula:
  * golden vectors: answers printed by the reference CPU programs compiled
    from LCS/LCS.cpp, SmithWaterman/SW.cpp, EditDistance/ED.cpp
    (tests/golden/ref_vectors.json, tests/golden/make_golden.py);
  * the reference's committed inputs LCS/Data (copied to tests/golden/data)
    and the published 21469 of LCS/GPU_Hyperlane_Share.txt:102;
  * brute force from the definitions on tiny inputs: LCS by subsequence
    enumeration, Smith-Waterman by enumerating every alignment of every pair
    of substrings, Levenshtein by breadth-first search over single edits;
  * the closed form ED.cpp(a, b) = min(lev(a, b), lev(a[1:], b)) that the
    row-0 boundary D(0,x) = x-1 implies (EditDistance/ED.cpp:48-54);
  * closed forms and invariants (identity, disjoint alphabets, symmetry).
"""
import itertools
import json
from collections import deque
from functools import lru_cache
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from gpumemefficient_amd.datafile import read_input

GOLD = Path(__file__).resolve().parent / "golden"
VEC = json.loads((GOLD / "ref_vectors.json").read_text())


# ---------------------------------------------------------------- brute force
def brute_lcs(a, b):
    a, b = list(a), list(b)
    best = 0
    for k in range(len(a), 0, -1):
        for idx in itertools.combinations(range(len(a)), k):
            sub = [a[i] for i in idx]
            it = iter(b)
            if all(any(c == d for d in it) for c in sub):
                return k
    return best


def brute_nw_alignments(a, b):
    """Scores of every global alignment of a and b (match +3, mismatch -3,
    gap -2), by explicit enumeration of alignment columns."""
    a, b = tuple(a), tuple(b)

    @lru_cache(maxsize=None)
    def rec(i, j):
        if i == len(a) and j == len(b):
            return frozenset([0])
        out = set()
        if i < len(a) and j < len(b):
            s = 3 if a[i] == b[j] else -3
            out |= {s + v for v in rec(i + 1, j + 1)}
        if i < len(a):
            out |= {-2 + v for v in rec(i + 1, j)}
        if j < len(b):
            out |= {-2 + v for v in rec(i, j + 1)}
        return frozenset(out)

    return rec(0, 0)


def brute_sw(a, b):
    """(last, max): max over all local alignments; last = best local
    alignment that ends at both ends (SW.cpp's bottom-right cell)."""
    n1, n2 = len(a), len(b)
    best, last = 0, 0
    for i in range(n1 + 1):
        for j in range(i, n1 + 1):
            for k in range(n2 + 1):
                for l in range(k, n2 + 1):
                    v = max(brute_nw_alignments(a[i:j], b[k:l]))
                    best = max(best, v)
                    if j == n1 and l == n2:
                        last = max(last, v)
    return last, best


def brute_lev(a, b, alphabet):
    """Minimum number of single-symbol insertions, deletions and
    substitutions turning a into b, by breadth-first search."""
    a, b = tuple(a), tuple(b)
    if a == b:
        return 0
    seen = {a}
    q = deque([(a, 0)])
    while q:
        s, d = q.popleft()
        nbrs = set()
        for i in range(len(s) + 1):
            for c in alphabet:
                nbrs.add(s[:i] + (c,) + s[i:])
        for i in range(len(s)):
            nbrs.add(s[:i] + s[i + 1:])
            for c in alphabet:
                nbrs.add(s[:i] + (c,) + s[i + 1:])
        for t in nbrs:
            if t == b:
                return d + 1
            if t not in seen and len(t) <= max(len(a), len(b)) + 1:
                seen.add(t)
                q.append((t, d + 1))
    raise AssertionError("unreachable")


# ------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("case", VEC["cases"], ids=lambda c: f"n{c['n']}a{c['alpha']}")
def test_golden_vectors(case):
    a, b = np.array(case["a"], np.int32), np.array(case["b"], np.int32)
    assert O.lcs(a, b) == case["lcs"]
    assert O.sw(a, b)[0] == case["sw_last"]
    assert O.ed(a, b) == case["ed"]


@pytest.mark.parametrize("name", sorted((GOLD / "data").glob("*.txt")), ids=lambda p: p.name)
def test_reference_data_files(name):
    import hashlib
    rec = VEC["reference_data"][name.name]
    assert hashlib.md5(name.read_bytes()).hexdigest() == rec["md5"]
    a, b = read_input(name)
    assert O.lcs(a, b) == rec["lcs"]
    assert O.sw(a, b)[0] == rec["sw_last"]
    assert O.ed(a, b) == rec["ed"]


def test_published_lcs_2_15():
    a, b = read_input(GOLD / "data" / "str1_2_15_str2_2_15.txt")
    assert O.lcs(a, b) == VEC["published"]["str1_2_15_str2_2_15.txt"]["lcs"] == 21469


def test_ed_quirk_kats():
    assert O.ed([1, 2], [2, 3]) == 1 and O.ed([1, 2], [2, 3], textbook=True) == 2
    assert O.ed([0, 1], [1, 0]) == 1 and O.ed([0, 1], [1, 0], textbook=True) == 2


# ---------------------------------------------------------------- brute force
def _tiny_pairs(seed, count, maxlen, alpha):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        n1, n2 = int(rng.integers(0, maxlen + 1)), int(rng.integers(0, maxlen + 1))
        yield rng.integers(0, alpha, n1).tolist(), rng.integers(0, alpha, n2).tolist()


def test_lcs_brute_force():
    for a, b in _tiny_pairs(1, 300, 8, 3):
        assert O.lcs(a, b) == brute_lcs(a, b), (a, b)


def test_sw_brute_force():
    for a, b in _tiny_pairs(2, 120, 5, 3):
        assert O.sw(a, b) == brute_sw(a, b), (a, b)


def test_levenshtein_brute_force_textbook_mode():
    for a, b in _tiny_pairs(3, 150, 4, 3):
        assert O.ed(a, b, textbook=True) == brute_lev(a, b, (0, 1, 2)), (a, b)


def test_ed_cpp_boundary_closed_form():
    """ED.cpp's D(0,x) = x-1 makes the first deletion of a[0] free:
    ED.cpp(a, b) = min(lev(a, b), lev(a[1:], b)) for n1 >= 1."""
    for a, b in _tiny_pairs(4, 150, 4, 3):
        if len(a) == 0:
            continue
        want = min(brute_lev(a, b, (0, 1, 2)), brute_lev(a[1:], b, (0, 1, 2)))
        assert O.ed(a, b) == want, (a, b)


# ------------------------------------------------- closed forms and invariants
def test_closed_forms():
    rng = np.random.default_rng(5)
    for n in (1, 2, 10, 100, 1000):
        a = rng.integers(0, 4, n, dtype=np.int32)
        assert O.lcs(a, a) == n
        assert O.ed(a, a) == 0 and O.ed(a, a, textbook=True) == 0
        assert O.sw(a, a) == (3 * n, 3 * n)
        c = a + 10  # disjoint alphabet
        assert O.lcs(a, c) == 0
        assert O.sw(a, c) == (0, 0)
        assert O.ed(a, c, textbook=True) == n


def test_symmetry_and_bounds():
    rng = np.random.default_rng(6)
    for _ in range(50):
        a = rng.integers(0, 4, int(rng.integers(1, 200)), dtype=np.int32)
        b = rng.integers(0, 4, int(rng.integers(1, 200)), dtype=np.int32)
        assert O.lcs(a, b) == O.lcs(b, a) <= min(a.size, b.size)
        assert O.sw(a, b)[1] == O.sw(b, a)[1]
        assert O.ed(a, b, True) == O.ed(b, a, True)
        assert abs(a.size - b.size) <= O.ed(a, b, True) <= max(a.size, b.size)


def test_table_and_last_row_agree():
    rng = np.random.default_rng(7)
    for prob in ("LCS", "ED", "SW"):
        for n1, n2 in [(1, 1), (5, 9), (37, 20), (64, 65)]:
            a = rng.integers(0, 4, n1, dtype=np.int32)
            b = rng.integers(0, 4, n2, dtype=np.int32)
            t = O.table(prob, a, b)
            assert t[-1, -1] == O.score(prob, a, b)[0]
            assert np.array_equal(O.last_row(prob, a, b), t[-1])
            if prob == "SW":
                assert t.max() == O.sw(a, b)[1]


def test_column_split_identity():
    """The column split the bit-parallel engine uses for single LCS/ED pairs:
    with F the last column of the table of a[:mid] against b (the pair's own
    row 0) and R the last column of the textbook table of the reversed
    a[mid:] against the reversed b, the score is min_y F(y) + R(n2 - y) for
    ED (ED.cpp's row 0 and the textbook one) and max_y for LCS."""
    rng = np.random.default_rng(41)
    for _ in range(150):
        n1 = int(rng.integers(2, 48))
        n2 = int(rng.integers(1, 70))
        alpha = int(rng.integers(2, 6))
        a = rng.integers(0, alpha, n1, dtype=np.int32)
        b = rng.integers(0, alpha, n2, dtype=np.int32)
        mid = int(rng.integers(1, n1))
        ra, rb = a[mid:][::-1].copy(), b[::-1].copy()
        for prob, textbook in (("ED", False), ("ED", True), ("LCS", False)):
            f = O.table(prob, a[:mid], b, textbook=textbook)[:, mid]
            r = O.table(prob, ra, rb, textbook=True)[:, n1 - mid]
            sums = [int(f[y]) + int(r[n2 - y]) for y in range(n2 + 1)]
            got = min(sums) if prob == "ED" else max(sums)
            assert got == O.score(prob, a, b, textbook=textbook)[0], (prob, textbook, n1, n2, mid)


def test_ed_rewrite_min_form_matches_reference_recurrence():
    """The kernels use D = min(up+1, left+1, diag+[a!=b]) in place of ED.cpp's
    `eq ? diag : 1+min(up,left,diag)` (EditDistance/ED.cpp:61-63), and
    LCS = max(up, left, diag+[a==b]) in place of LCS.cpp:58-63.  Check the
    rewrites cell by cell on full tables."""
    rng = np.random.default_rng(8)
    for textbook in (False, True):
        for _ in range(40):
            n1, n2 = int(rng.integers(1, 40)), int(rng.integers(1, 40))
            a = rng.integers(0, 3, n1)
            b = rng.integers(0, 3, n2)
            t = O.table("ED", a, b, textbook)
            for y in range(1, n2 + 1):
                for x in range(1, n1 + 1):
                    ne = int(a[x - 1] != b[y - 1])
                    assert t[y, x] == min(t[y - 1, x] + 1, t[y, x - 1] + 1, t[y - 1, x - 1] + ne)
            l = O.table("LCS", a, b)
            for y in range(1, n2 + 1):
                for x in range(1, n1 + 1):
                    eq = int(a[x - 1] == b[y - 1])
                    assert l[y, x] == max(l[y - 1, x], l[y, x - 1], l[y - 1, x - 1] + eq)


def test_planted_mistake_is_caught():
    """A plausible slip (ED.cpp boundary replaced by the textbook one) must
    fail the golden vectors: the quirk cases differ."""
    quirk = [c for c in VEC["cases"] if "quirk" in c.get("note", "")]
    assert quirk
    assert any(O.ed(c["a"], c["b"], textbook=True) != c["ed"] for c in quirk)


# ----------------------------------------------------------------------------
# In-place table stencils (SAT/GPU_Hyperlane_Share.cu:169,
# 2D-SOR/nested_loop/GPU_Hyperlane_Share.cu:169): pinned by a closed form
# (SAT), hand-computed known answers and traversal-order invariance (SOR).
# ----------------------------------------------------------------------------

def _st_table(n1, n2, seed, lo=0, hi=10000):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, size=(n2 + 2, n1 + 3), dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("n1,n2,lo,hi", [(1, 1, 0, 10000), (7, 5, 0, 10000), (60, 45, 0, 10000),
                                         (33, 70, -(2 ** 31), 2 ** 31)])
def test_sat_closed_form(n1, n2, lo, hi):
    """T(y,x) = sum of the interior inputs above-left of (y,x) plus the
    boundary: T(y,0) + T(0,x) - T(0,0), all mod 2^32 -- independent of the
    recurrence (a wrong sign of the diagonal, or an old/new mix-up, breaks it)."""
    t = _st_table(n1, n2, n1 * 31 + n2, lo, hi)
    got = O.stencil("SAT", t)
    v = t.astype(np.int64)
    inner = np.zeros((n2 + 1, n1 + 1), np.int64)
    inner[1:, 1:] = v[1:n2 + 1, 1:n1 + 1].cumsum(0).cumsum(1)
    want = inner + v[: n2 + 1, :1] + v[:1, : n1 + 1] - v[0, 0]
    want = ((want + 2 ** 31) % 2 ** 32 - 2 ** 31).astype(np.int32)
    assert np.array_equal(got[: n2 + 1, : n1 + 1], want)
    # Boundary row n2+1 and columns n1+1, n1+2 are untouched.
    assert np.array_equal(got[n2 + 1], t[n2 + 1])
    assert np.array_equal(got[:, n1 + 1:], t[:, n1 + 1:])


def test_sor_known_answers():
    """A 2 x 2 interior worked by hand: (1,1) sees the old right and down
    values, (1,2) the new left, (2,1) the new up, (2,2) both new."""
    t = np.array([[0, 10, 20, 0, 0],
                  [30, 40, 50, 60, 0],
                  [70, 80, 90, 100, 0],
                  [0, 110, 120, 0, 0]], dtype=np.int32)
    got = O.stencil("SOR", t)
    a = (30 + 10 + 40 + 50 + 80) // 5          # 42
    b = (a + 20 + 50 + 60 + 90) // 5            # 52
    c = (70 + a + 80 + 90 + 110) // 5           # 78
    d = (c + b + 90 + 100 + 120) // 5           # 88
    assert (a, b, c, d) == (42, 52, 78, 88)
    assert got[1, 1:3].tolist() == [a, b] and got[2, 1:3].tolist() == [c, d]
    assert np.array_equal(got[0], t[0]) and np.array_equal(got[3], t[3])


def test_sor_truncates_toward_zero():
    t = np.zeros((3, 4), np.int32)
    t[1, 1] = -7  # sum -7: C division gives -1, floor division would give -2
    assert O.stencil("SOR", t)[1, 1] == -1
    t[1, 1] = 7
    assert O.stencil("SOR", t)[1, 1] == 1


@pytest.mark.parametrize("c", [0, 1, -3, 9999])
def test_sor_constant_table_is_fixed(c):
    t = np.full((12, 15), c, np.int32)
    assert np.array_equal(O.stencil("SOR", t), t)


def _sor_antidiagonal(t):
    """The same sweep in anti-diagonal order (the hyperlane tiles' order):
    it must equal the row-major one."""
    t = t.astype(np.int64).copy()
    n2, n1 = t.shape[0] - 2, t.shape[1] - 3
    for d in range(2, n1 + n2 + 1):
        for y in range(max(1, d - n1), min(n2, d - 1) + 1):
            x = d - y
            s = t[y, x - 1] + t[y - 1, x] + t[y, x] + t[y, x + 1] + t[y + 1, x]
            t[y, x] = int(s / 5)  # values stay small: no wrap, truncation toward zero
    return t.astype(np.int32)


@pytest.mark.parametrize("n1,n2", [(1, 1), (6, 9), (17, 11)])
def test_sor_traversal_order_invariant(n1, n2):
    t = _st_table(n1, n2, n1 + 100 * n2, -5000, 5000)
    assert np.array_equal(O.stencil("SOR", t), _sor_antidiagonal(t))
