"""wf_align_noalloc (no device allocation, static per-GPU scratch) against
wf_align on the same inputs, through the C ABI: two calls in a row, the
second smaller, for each problem and engine family, then a pair too large
for the scratch area.  Integer work: bit-exact."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib_and_ctx():
    from gpumemefficient_amd import _lib
    lib = _lib.load()
    ctx = C.c_void_p()
    st = lib.wf_ctx_create(0, C.byref(ctx))
    if st == _lib.WF_ERR_DEVICE:
        pytest.skip("no HIP device on this host")
    assert st == _lib.WF_OK, st
    return _lib, lib, ctx


def _call(fn, ctx, prob, a, b, tile=None):
    from gpumemefficient_amd import _lib
    r = _lib.wf_result()
    pa = a.ctypes.data_as(C.POINTER(C.c_int32))
    pb = b.ctypes.data_as(C.POINTER(C.c_int32))
    st = fn(ctx, _lib.PROBLEMS[prob], pa, C.c_int64(a.size), pb, C.c_int64(b.size),
            C.byref(tile) if tile is not None else None, C.byref(r))
    return st, (r.last, r.max_score, r.status, r.cells)


# (problem, alphabet, first pair, second pair): LCS and ED of at most 32
# symbols take the bit-parallel engine (the first pair is wide enough for
# its column split), SW the asynchronous cell engine, ED over 40 symbols the
# cell engine with compared codes.
CASES = [
    ("SW", 4, (3000, 2100), (700, 513)),
    ("LCS", 4, (4000, 3000), (1500, 900)),
    ("ED", 20, (2600, 2500), (1100, 640)),
    ("ED", 40, (2000, 1700), (333, 1000)),
    ("SW", 8, (1800, 1800), (64, 65)),
]


@pytest.mark.parametrize("prob,alpha,first,second", CASES)
def test_noalloc_matches_wf_align_twice(prob, alpha, first, second):
    _lib, lib, ctx = _lib_and_ctx()
    try:
        rng = np.random.default_rng(131 * alpha + len(prob))
        for n1, n2 in (first, second):
            a = rng.integers(0, alpha, n1, dtype=np.int32)
            b = rng.integers(0, alpha, n2, dtype=np.int32)
            print(f"  {prob} alpha={alpha} n1={n1} n2={n2}", flush=True)
            st_ref, ref = _call(lib.wf_align, ctx, prob, a, b)
            st, got = _call(lib.wf_align_noalloc, ctx, prob, a, b)
            assert st_ref == _lib.WF_OK and st == _lib.WF_OK, (st_ref, st)
            assert got == ref, (got, ref)
    finally:
        lib.wf_ctx_destroy(ctx)


def test_noalloc_rejects_pairs_beyond_the_scratch_and_keeps_working():
    _lib, lib, ctx = _lib_and_ctx()
    try:
        rng = np.random.default_rng(5)
        # SW 32768 x 32768 on 512-row strips: 63 boundary rows of 32768
        # words, 16.5 MB -- fits; 65536 x 65536: 127 rows, 66.6 MB -- does not.
        big = rng.integers(0, 4, 65536, dtype=np.int32)
        st, _ = _call(lib.wf_align_noalloc, ctx, "SW", big, big)
        assert st == _lib.WF_ERR_SCRATCH, st
        assert "scratch" in lib.wf_strerror(st).decode()
        a = rng.integers(0, 4, 32768, dtype=np.int32)
        b = rng.integers(0, 4, 8192, dtype=np.int32)
        st_ref, ref = _call(lib.wf_align, ctx, "SW", a, b)
        st, got = _call(lib.wf_align_noalloc, ctx, "SW", a, b)
        assert st_ref == _lib.WF_OK and st == _lib.WF_OK and got == ref, (st, got, ref)
    finally:
        lib.wf_ctx_destroy(ctx)


def test_noalloc_runs_an_sw_pair_unsplit_where_its_split_does_not_fit():
    """SW 65536 x 51200, bottom-right cell only: 100 strips of 512 rows, 99
    boundary rows of 65536 words (51.9 MB) fit the scratch area; the column
    split that wf_align plans (three pairs, 300 strips) would need 1.5 times
    that.  wf_align_noalloc runs it unsplit, with the same result."""
    _lib, lib, ctx = _lib_and_ctx()
    try:
        rng = np.random.default_rng(6)
        a = rng.integers(0, 4, 65536, dtype=np.int32)
        b = rng.integers(0, 4, 51200, dtype=np.int32)
        tile = _lib.wf_tile(0, 0, 0, _lib.WF_TILE_NO_MAX)
        st_ref, ref = _call(lib.wf_align, ctx, "SW", a, b, tile)
        st, got = _call(lib.wf_align_noalloc, ctx, "SW", a, b, tile)
        assert st_ref == _lib.WF_OK and st == _lib.WF_OK and got == ref, (st_ref, st, got, ref)
        assert ref[1] == -1
    finally:
        lib.wf_ctx_destroy(ctx)
