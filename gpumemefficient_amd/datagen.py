"""Seeded synthetic workloads.

Every generator uses numpy's PCG64 (``numpy.random.default_rng(seed)``), so a
(seed, size) pair always gives the same symbols.  The reference's
generator.cpp seeds ``rand()`` with the wall clock (LCS/generator.cpp:36) and
is not reproducible; these keep its alphabet convention (symbols 0..k-1) and
its file format (gpumemefficient_amd.datafile).

Workloads:
  lcs_pair      LCS, two i.i.d. uniform 4-letter sequences (8192, seed 1)
  ed_pair       edit distance, i.i.d. uniform 20-letter (65536, seed 2)
  sw_pair       Smith-Waterman, 4-letter (131072, seed 3) with a planted
                4096-long homologous segment: a copy goes into a at a random
                offset, a copy with 10% of positions substituted by a
                different letter goes into b at another random offset
  batch_pairs   4096 pairs, both lengths independently log-uniform in
                [64, 4096], 4-letter (seed 4)
  serving_requests  Poisson request stream, 50% ED (20-letter) / 50% SW
                (4-letter); lengths 80% U[100,1000], 18% U[1000,16384],
                2% U[16384,131072] (seed 5)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def uniform_pair(n1: int, n2: int, alphabet: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    a = rng.integers(0, alphabet, n1, dtype=np.int32)
    b = rng.integers(0, alphabet, n2, dtype=np.int32)
    return a, b


def lcs_pair(n: int = 8192, seed: int = 1):
    return uniform_pair(n, n, 4, seed)


def ed_pair(n: int = 65536, seed: int = 2):
    return uniform_pair(n, n, 20, seed)


def sw_pair(n: int = 131072, seed: int = 3, seg: int = 4096, sub_rate: float = 0.10):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 4, n, dtype=np.int32)
    b = rng.integers(0, 4, n, dtype=np.int32)
    seg = min(seg, n)
    s = rng.integers(0, 4, seg, dtype=np.int32)
    oa = int(rng.integers(0, n - seg + 1))
    ob = int(rng.integers(0, n - seg + 1))
    t = s.copy()
    mask = rng.random(seg) < sub_rate
    t[mask] = (t[mask] + rng.integers(1, 4, int(mask.sum()), dtype=np.int32)) % 4
    a[oa:oa + seg] = s
    b[ob:ob + seg] = t
    return a, b


def loguniform_lengths(rng: np.random.Generator, count: int, lo: int, hi: int) -> np.ndarray:
    u = rng.uniform(np.log(lo), np.log(hi), count)
    return np.clip(np.floor(np.exp(u) + 0.5).astype(np.int64), lo, hi)


def batch_pairs(count: int = 4096, lo: int = 64, hi: int = 4096, seed: int = 4, alphabet: int = 4):
    """List of (a, b) with independently log-uniform lengths."""
    rng = np.random.default_rng(seed)
    n1 = loguniform_lengths(rng, count, lo, hi)
    n2 = loguniform_lengths(rng, count, lo, hi)
    return [(rng.integers(0, alphabet, int(x), dtype=np.int32), rng.integers(0, alphabet, int(y), dtype=np.int32))
            for x, y in zip(n1, n2)]


@dataclass
class Request:
    rid: int
    arrival_s: float     # scheduled arrival, seconds from stream start
    problem: str         # "ED" or "SW"
    a: np.ndarray
    b: np.ndarray

    @property
    def cells(self) -> int:
        return int(self.a.size) * int(self.b.size)


BUCKETS = ((0.80, 100, 1000), (0.18, 1000, 16384), (0.02, 16384, 131072))


def serving_requests(rate_per_s: float, count: int, seed: int = 5) -> list[Request]:
    """Poisson arrivals at `rate_per_s`; each request picks ED or SW with
    probability 1/2 and a length bucket, then both lengths uniformly in it."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, count)
    t = np.cumsum(gaps)
    probs = np.array([b[0] for b in BUCKETS])
    out = []
    for i in range(count):
        k = int(rng.choice(len(BUCKETS), p=probs))
        _, lo, hi = BUCKETS[k]
        n1, n2 = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        prob = "ED" if rng.random() < 0.5 else "SW"
        alpha = 20 if prob == "ED" else 4
        out.append(Request(i, float(t[i]), prob, rng.integers(0, alpha, n1, dtype=np.int32),
                           rng.integers(0, alpha, n2, dtype=np.int32)))
    return out
