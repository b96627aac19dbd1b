"""Python host interface over libwfdp.so.

Two layers:

* the reference's operator API, one function per problem with the
  reference's argument meaning -- ``LCS(n1, n2, arr1, arr2, paddX, paddY,
  table)`` (LCS/GPU_Hyperlane_Share.h:5), ``SW(..., tX, tY)``
  (SmithWaterman/GPU_Hyperlane_Share.h:5), ``ED(..., tX, tY)``
  (EditDistance/GPU_Hyperlane_Share.h:5) -- returning the bottom-right cell;
* contexts, batches and plans for the batched and served workloads.

Every call goes through the HIP library; nothing here computes a DP cell.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import WfError, check

PROBLEM_IDS = {"LCS": _lib.WF_LCS, "ED": _lib.WF_ED, "SW": _lib.WF_SW}


def _problem_id(problem) -> int:
    if isinstance(problem, str):
        return PROBLEM_IDS[problem.upper()]
    p = int(problem)
    if p not in (0, 1, 2):
        raise ValueError(f"unknown problem {problem!r}")
    return p


def _stencil_id(op) -> int:
    if isinstance(op, str):
        try:
            return _lib.STENCILS[op.upper()]
        except KeyError:
            raise ValueError(f"unknown stencil {op!r}; expected one of {sorted(_lib.STENCILS)}") from None
    return int(op)


def _as_i32(x) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))
    return arr


def _ptr(arr: np.ndarray):
    return arr.ctypes.data_as(C.POINTER(C.c_int32))


@dataclass
class Result:
    last: int        # bottom-right cell: what the reference programs print
    max_score: int   # SW maximum cell (local alignment score); 0 for LCS/ED
    status: int
    kernel_ms: float
    cells: int

    @classmethod
    def _from(cls, r: _lib.wf_result) -> "Result":
        return cls(int(r.last), int(r.max_score), int(r.status), float(r.kernel_ms), int(r.cells))


@dataclass
class Tile:
    """Tile request: tile_w = steps between boundary exchanges, strip_rows =
    rows swept by one workgroup, rows_per_lane = DP rows per SIMD lane.
    Zero means "choose"."""
    tile_w: int = 0
    strip_rows: int = 0
    rows_per_lane: int = 0
    ed_textbook: bool = False
    cache: bool = False
    engine: str = ""      # "lockstep", "async", "bitpar", "bitpar_staged" or "" (default)
    hint: bool = False    # the fields are hints (WF_TILE_HINT): the planner picks the variant
    no_max: bool = False  # SW: bottom-right cell only, no running maximum (WF_TILE_NO_MAX; max_score -1)

    def _c(self) -> _lib.wf_tile:
        flags = ((_lib.WF_TILE_ED_TEXTBOOK if self.ed_textbook else 0) | (_lib.WF_TILE_CACHE if self.cache else 0)
                 | (_lib.WF_TILE_LOCKSTEP if self.engine == "lockstep" else 0)
                 | (_lib.WF_TILE_ASYNC if self.engine == "async" else 0)
                 | (_lib.WF_TILE_BITPAR if self.engine == "bitpar" else 0)
                 | (_lib.WF_TILE_BITPAR_STAGED if self.engine == "bitpar_staged" else 0)
                 | (_lib.WF_TILE_HINT if self.hint else 0)
                 | (_lib.WF_TILE_NO_MAX if self.no_max else 0))
        return _lib.wf_tile(self.tile_w, self.strip_rows, self.rows_per_lane, flags)

    @classmethod
    def _from(cls, t: _lib.wf_tile) -> "Tile":
        eng = ("lockstep" if t.flags & _lib.WF_TILE_LOCKSTEP else "async" if t.flags & _lib.WF_TILE_ASYNC
               else "bitpar" if t.flags & _lib.WF_TILE_BITPAR
               else "bitpar_staged" if t.flags & _lib.WF_TILE_BITPAR_STAGED else "")
        return cls(t.tile_w, t.strip_rows, t.rows_per_lane, bool(t.flags & _lib.WF_TILE_ED_TEXTBOOK),
                   bool(t.flags & _lib.WF_TILE_CACHE), eng, bool(t.flags & _lib.WF_TILE_HINT),
                   bool(t.flags & _lib.WF_TILE_NO_MAX))


@dataclass
class Job:
    problem: int | str
    a: Sequence[int] | np.ndarray
    b: Sequence[int] | np.ndarray
    ed_textbook: bool = False
    no_max: bool = False  # SW: bottom-right cell only (WF_TILE_NO_MAX as a job flag)


class _JobArray:
    """Keeps the int32 arrays alive while C holds pointers into them."""

    def __init__(self, jobs: Iterable[Job]):
        self.arrays = []
        jl = list(jobs)
        self.c = (_lib.wf_job * max(1, len(jl)))()
        for i, j in enumerate(jl):
            a, b = _as_i32(j.a), _as_i32(j.b)
            self.arrays += [a, b]
            flags = (_lib.WF_TILE_ED_TEXTBOOK if j.ed_textbook else 0) | (_lib.WF_TILE_NO_MAX if j.no_max else 0)
            self.c[i] = _lib.wf_job(_problem_id(j.problem), flags, _ptr(a), a.size, _ptr(b), b.size)
        self.n = len(jl)


class Plan:
    """Encoded pairs resident in HBM; ``run`` re-executes the DP."""

    def __init__(self, ctx: "Context", jobs: Iterable[Job], tile: Tile | None = None):
        self.ctx = ctx
        self._jobs = _JobArray(jobs)
        self._h = C.c_void_p()
        t = tile._c() if tile else None
        check(ctx._lib.wf_plan_create(ctx._h, self._jobs.c, self._jobs.n, C.byref(t) if t else None,
                                      C.byref(self._h)), "wf_plan_create")
        self.n = self._jobs.n

    def run(self, per_pair: bool = False) -> float:
        check(self.ctx._lib.wf_plan_run(self._h, _lib.WF_RUN_PER_PAIR if per_pair else _lib.WF_RUN_PERSISTENT),
              "wf_plan_run")
        return self.kernel_ms

    def capture(self, per_pair: bool = False) -> None:
        """Capture this plan's launches into a hipGraph; later run() calls of
        the same mode replay it."""
        check(self.ctx._lib.wf_plan_capture(self._h, _lib.WF_RUN_PER_PAIR if per_pair else _lib.WF_RUN_PERSISTENT),
              "wf_plan_capture")

    @property
    def kernel_ms(self) -> float:
        return float(self.ctx._lib.wf_plan_kernel_ms(self._h))

    def results(self, raise_on_error: bool = True) -> list[Result]:
        out = (_lib.wf_result * max(1, self.n))()
        st = self.ctx._lib.wf_plan_results(self._h, out, self.n)
        if raise_on_error:
            check(st, "wf_plan_results")
        return [Result._from(out[i]) for i in range(self.n)]

    def describe(self) -> dict:
        e, r, w, k, g = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        nt = C.c_int64()
        check(self.ctx._lib.wf_plan_describe(self._h, C.byref(e), C.byref(r), C.byref(w), C.byref(k), C.byref(nt),
                                             C.byref(g)))
        return {"engine": _lib.ENGINES[e.value], "rows_per_lane": r.value, "waves": w.value,
                "chunk": k.value, "tasks": nt.value, "grid": g.value}

    def close(self) -> None:
        if self._h:
            self.ctx._lib.wf_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device context (stream, arenas, events)."""

    def __init__(self, device: int = 0, timeout_s: float | None = None):
        self._lib = _lib.load()
        self._h = C.c_void_p()
        check(self._lib.wf_ctx_create(device, C.byref(self._h)), "wf_ctx_create")
        if timeout_s is not None:
            check(self._lib.wf_ctx_set_timeout(self._h, float(timeout_s)))

    def info(self) -> dict:
        d = _lib.wf_device_info()
        check(self._lib.wf_device_query(self._h, C.byref(d)))
        return {"compute_units": d.compute_units, "max_clock_mhz": d.max_clock_mhz,
                "simds_per_cu": d.simds_per_cu, "wave_size": d.wave_size, "total_mem": d.total_mem,
                "free_mem": d.free_mem, "arch": d.arch.decode()}

    def align(self, problem, a, b, tile: Tile | None = None, ed_textbook: bool = False) -> Result:
        a, b = _as_i32(a), _as_i32(b)
        t = tile._c() if tile else _lib.wf_tile(0, 0, 0, 0)
        if ed_textbook:
            t.flags |= _lib.WF_TILE_ED_TEXTBOOK
        out = _lib.wf_result()
        check(self._lib.wf_align(self._h, _problem_id(problem), _ptr(a), a.size, _ptr(b), b.size, C.byref(t),
                                 C.byref(out)), "wf_align")
        return Result._from(out)

    def align_table(self, problem, a, b, ed_textbook: bool = False) -> tuple[Result, np.ndarray]:
        """Score of one pair plus its whole DP table as an (n2, n1) int32
        array, table[y-1, x-1] = H(y, x) (the reference drivers' DEBUG dump)."""
        a, b = _as_i32(a), _as_i32(b)
        t = _lib.wf_tile(0, 0, 0, _lib.WF_TILE_ED_TEXTBOOK if ed_textbook else 0)
        table = np.empty((b.size, a.size), dtype=np.int32)
        out = _lib.wf_result()
        check(self._lib.wf_align_table(self._h, _problem_id(problem), _ptr(a), a.size, _ptr(b), b.size, C.byref(t),
                                       _ptr(table), C.byref(out)), "wf_align_table")
        return Result._from(out), table

    def align_naive(self, problem, a, b, method: str = "diagonal", ed_textbook: bool = False,
                    want_table: bool = False) -> tuple[Result, np.ndarray | None]:
        """The reference's naive baselines: "diagonal" (one launch per cell
        anti-diagonal) or "tile" (one launch per anti-diagonal of 64 x 128
        tiles).  With want_table the (n2, n1) table comes back as well."""
        a, b = _as_i32(a), _as_i32(b)
        m = {"diagonal": _lib.WF_NAIVE_DIAGONAL, "tile": _lib.WF_NAIVE_TILE}[method]
        t = _lib.wf_tile(0, 0, 0, _lib.WF_TILE_ED_TEXTBOOK if ed_textbook else 0)
        table = np.empty((b.size, a.size), dtype=np.int32) if want_table else None
        out = _lib.wf_result()
        check(self._lib.wf_align_naive(self._h, _problem_id(problem), _ptr(a), a.size, _ptr(b), b.size, C.byref(t),
                                       m, _ptr(table) if table is not None else None, C.byref(out)),
              "wf_align_naive")
        return Result._from(out), table

    def stencil(self, op, table, strip_rows: int = 0) -> tuple[Result, np.ndarray]:
        """One in-place sweep of the reference's SAT or 2D-SOR stencil ("SAT"
        / "SOR") over an (n2 + 2, n1 + 3) int32 table (boundary row 0, row
        n2 + 1, columns 0, n1 + 1, n1 + 2).  Returns the result and the swept
        copy; `table` is not modified."""
        t = np.array(table, dtype=np.int32, copy=True, order="C")
        if t.ndim != 2 or t.shape[0] < 2 or t.shape[1] < 3:
            raise ValueError("table must be (n2 + 2, n1 + 3) with n1, n2 >= 0")
        n2, n1 = t.shape[0] - 2, t.shape[1] - 3
        tl = _lib.wf_tile(0, int(strip_rows), 0, 0)
        out = _lib.wf_result()
        check(self._lib.wf_stencil(self._h, _stencil_id(op), n1, n2, _ptr(t), t.shape[1], C.byref(tl),
                                   C.byref(out)), "wf_stencil")
        return Result._from(out), t

    def stencil_device(self, op, dev_ptr: int, n1: int, n2: int, ld: int, strip_rows: int = 0) -> Result:
        """The same sweep on a table already in HBM (a device pointer, e.g. a
        torch tensor's data_ptr(); pitch ld % 4 == 0, 16-byte aligned)."""
        tl = _lib.wf_tile(0, int(strip_rows), 0, 0)
        out = _lib.wf_result()
        check(self._lib.wf_stencil_device(self._h, _stencil_id(op), int(n1), int(n2), C.c_void_p(int(dev_ptr)),
                                          int(ld), C.byref(tl), C.byref(out)), "wf_stencil_device")
        return Result._from(out)

    def align_batch(self, jobs: Iterable[Job], tile: Tile | None = None) -> list[Result]:
        ja = _JobArray(jobs)
        out = (_lib.wf_result * max(1, ja.n))()
        t = tile._c() if tile else None
        check(self._lib.wf_align_batch(self._h, ja.c, ja.n, C.byref(t) if t else None, out), "wf_align_batch")
        return [Result._from(out[i]) for i in range(ja.n)]

    def plan(self, jobs: Iterable[Job], tile: Tile | None = None) -> Plan:
        return Plan(self, jobs, tile)

    def autotune(self, jobs: Iterable[Job], reps: int = 2) -> tuple[Tile, float]:
        """Time every compiled variant on `jobs`; returns the fastest tile and
        its kernel time.  Later plans of the same workload class without an
        explicit tile use it."""
        ja = _JobArray(jobs)
        t = _lib.wf_tile()
        ms = C.c_double()
        check(self._lib.wf_autotune(self._h, ja.c, ja.n, reps, C.byref(t), C.byref(ms)), "wf_autotune")
        return Tile._from(t), ms.value

    def close(self) -> None:
        if self._h:
            self._lib.wf_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def variants() -> list[dict]:
    """Compiled kernel variants: engine, rows_per_lane, waves, chunk."""
    lib = _lib.load()
    out = []
    for i in range(lib.wf_variant_count()):
        e, r, w, k = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.wf_variant_info(i, C.byref(e), C.byref(r), C.byref(w), C.byref(k)))
        out.append({"engine": _lib.ENGINES[e.value], "rows_per_lane": r.value, "waves": w.value,
                    "chunk": k.value})
    return out


def tile_for(v: dict, **kw) -> Tile:
    """The Tile that selects exactly variant `v` (see variants())."""
    return Tile(tile_w=v["chunk"], strip_rows=64 * v["waves"] * v["rows_per_lane"],
                rows_per_lane=v["rows_per_lane"], engine=v["engine"], **kw)


def tune_load(path) -> None:
    check(_lib.load().wf_tune_load(str(path).encode()), f"wf_tune_load({path})")


def tune_save(path) -> None:
    check(_lib.load().wf_tune_save(str(path).encode()), f"wf_tune_save({path})")


def tune_clear() -> None:
    _lib.load().wf_tune_clear()


def tuned(jobs: Iterable[Job], ctx: "Context | None" = None) -> Tile | None:
    """The tile the autotune table holds for `jobs` on this device (what a
    plan without an explicit tile uses), or None."""
    c = ctx or default_context()
    ja = _JobArray(jobs)
    t = _lib.wf_tile()
    st = c._lib.wf_tune_lookup_tile(c._h, ja.c, ja.n, C.byref(t))
    if st == _lib.WF_ERR_NOKERNEL:
        return None
    check(st, "wf_tune_lookup_tile")
    return Tile._from(t)


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def _tile_from_ref(tX: int | None, tY: int | None, cache: bool = False) -> Tile:
    # tY rows per strip and tX columns per tile, as on the reference's argv
    # (SmithWaterman/GPU_Hyperlane_Share.cpp:56-59), passed as hints: the
    # planner runs its fastest variant (every variant gives the same score).
    return Tile(tile_w=int(tX or 0), strip_rows=int(tY or 0), cache=cache, hint=True)


def LCS(n1: int, n2: int, arr1, arr2, paddX: int = 3, paddY: int = 1, table=None,
        tX: int | None = 32, tY: int | None = 256, cache: bool = False) -> int:
    """Reference signature LCS(n1, n2, arr1, arr2, paddX, paddY, table)
    (LCS/GPU_Hyperlane_Share.h:5); tile defaults to the hard-coded 32x256 of
    LCS/GPU_Hyperlane_Share.cu:703-704.  Returns the LCS length."""
    del paddX, paddY, table  # host table is never written by the reference outside DEBUG builds
    a, b = _as_i32(arr1)[:n1], _as_i32(arr2)[:n2]
    return default_context().align("LCS", a, b, _tile_from_ref(tX, tY, cache)).last


def SW(n1: int, n2: int, arr1, arr2, paddX: int = 3, paddY: int = 1, table=None,
       tX: int | None = None, tY: int | None = None, cache: bool = False) -> int:
    """Reference signature SW(n1, n2, arr1, arr2, paddX, paddY, table, tX, tY)
    (SmithWaterman/GPU_Hyperlane_Share.h:5).  Returns the bottom-right cell
    ("last", SmithWaterman/SW.cpp:80); the running maximum, which the
    reference does not compute, is skipped (no_max)."""
    del paddX, paddY, table
    a, b = _as_i32(arr1)[:n1], _as_i32(arr2)[:n2]
    t = _tile_from_ref(tX, tY, cache)
    t.no_max = True
    return default_context().align("SW", a, b, t).last


def ED(n1: int, n2: int, arr1, arr2, paddX: int = 3, paddY: int = 1, table=None,
       tX: int | None = None, tY: int | None = None, cache: bool = False, textbook: bool = False) -> int:
    """Reference signature ED(n1, n2, arr1, arr2, paddX, paddY, table, tX, tY)
    (EditDistance/GPU_Hyperlane_Share.h:5).  Boundary of EditDistance/ED.cpp
    (D(0,x) = x-1) unless textbook=True (the reference GPU's D(0,x) = x)."""
    del paddX, paddY, table
    a, b = _as_i32(arr1)[:n1], _as_i32(arr2)[:n2]
    t = _tile_from_ref(tX, tY, cache)
    t.ed_textbook = textbook
    return default_context().align("ED", a, b, t).last


def SOR(n1: int, n2: int, table, tX: int | None = None, tY: int | None = None, kind: str = "SOR") -> np.ndarray:
    """Reference signature SOR(n1, n2, table[, tX, tY]) of the hyperlane SAT
    and 2D-SOR programs (SAT/GPU_Hyperlane_Share.h:5,
    2D-SOR/nested_loop/GPU_Hyperlane_Share.h:5), whose host function is named
    SOR in both; kind="SAT" selects the summed-area table.  `table` holds at
    least (n2 + 2) rows of n1 + 3 values (the reference reads n2 + 3 rows,
    SAT/GPU_Hyperlane_Share.cpp:23-29, and uploads n2 + 2).  The reference
    leaves `table` unchanged (its copy back is commented out,
    SAT/GPU_Hyperlane_Share.cu:317); this returns the swept (n2 + 2, n1 + 3)
    table instead.  tX and tY only shape the reference's tiles."""
    del tX, tY
    t = np.asarray(table, dtype=np.int32).reshape(-1)[: (n2 + 2) * (n1 + 3)].reshape(n2 + 2, n1 + 3)
    return default_context().stencil(kind, t)[1]


__all__ = ["Context", "Plan", "Job", "Tile", "Result", "WfError", "LCS", "SW", "ED", "SOR", "default_context",
           "variants", "tile_for", "tuned"]
