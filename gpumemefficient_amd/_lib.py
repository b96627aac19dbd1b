"""ctypes binding of libwfdp.so (include/wfdp.h).

The library is the product path: if it is missing this module raises, it never
falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("WFDP_LIB", PKG / "lib" / "libwfdp.so"))

WF_LCS, WF_ED, WF_SW = 0, 1, 2
PROBLEMS = {"LCS": WF_LCS, "ED": WF_ED, "SW": WF_SW}

WF_OK = 0
WF_ERR_INVALID = 1
WF_ERR_DEVICE = 2
WF_ERR_NOMEM = 3
WF_ERR_ALPHABET = 4
WF_ERR_TIMEOUT = 5
WF_ERR_NOKERNEL = 6
WF_ERR_AGAIN = 7
WF_ERR_SCRATCH = 8
WF_NOALLOC_SCRATCH_BYTES = 64 << 20

WF_TILE_ED_TEXTBOOK = 0x1
WF_TILE_CACHE = 0x2
WF_TILE_LOCKSTEP = 0x4
WF_TILE_ASYNC = 0x8
WF_TILE_NO_COPYBACK = 0x10
WF_TILE_BITPAR = 0x20
WF_TILE_BITPAR_STAGED = 0x40
WF_TILE_HINT = 0x80
WF_TILE_NO_MAX = 0x100
ENGINES = ("lockstep", "async", "bitpar", "bitpar_staged")

WF_TABLE_MAX_CELLS = 1 << 28
WF_NAIVE_DIAGONAL = 0
WF_NAIVE_TILE = 1

WF_SAT, WF_SOR = 0, 1
STENCILS = {"SAT": WF_SAT, "SOR": WF_SOR}

WF_RUN_PERSISTENT = 0
WF_RUN_PER_PAIR = 1

# Every symbol include/wfdp.h declares (checked by the CPU test suite).
EXPORTS = [
    "wf_ctx_create", "wf_ctx_destroy", "wf_device_query", "wf_strerror", "wf_abi_version",
    "wf_ctx_set_timeout", "wf_align", "wf_align_table", "wf_align_naive", "wf_stencil", "wf_stencil_device",
    "wf_align_batch", "wf_plan_create", "wf_plan_run",
    "wf_plan_results", "wf_plan_kernel_ms", "wf_plan_describe", "wf_plan_destroy",
    "wf_variant_count", "wf_variant_info", "wf_plan_capture", "wf_autotune", "wf_tune_load", "wf_tune_save",
    "wf_tune_clear", "wf_server_create", "wf_server_submit", "wf_server_poll", "wf_server_wait_idle",
    "wf_server_stats", "wf_server_replay", "wf_server_destroy", "wf_align_noalloc", "wf_tune_lookup_tile",
]


class wf_tile(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("strip_rows", C.c_int32), ("rows_per_lane", C.c_int32),
                ("flags", C.c_int32)]


class wf_result(C.Structure):
    _fields_ = [("last", C.c_int32), ("max_score", C.c_int32), ("status", C.c_int32),
                ("reserved", C.c_int32), ("kernel_ms", C.c_double), ("cells", C.c_uint64)]


class wf_job(C.Structure):
    _fields_ = [("problem", C.c_int32), ("flags", C.c_int32), ("a", C.POINTER(C.c_int32)),
                ("n1", C.c_int64), ("b", C.POINTER(C.c_int32)), ("n2", C.c_int64)]


class wf_device_info(C.Structure):
    _fields_ = [("compute_units", C.c_int32), ("max_clock_mhz", C.c_int32), ("simds_per_cu", C.c_int32),
                ("wave_size", C.c_int32), ("total_mem", C.c_int64), ("free_mem", C.c_int64),
                ("arch", C.c_char * 32)]


class WfError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = strerror(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load libwfdp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gpumemefficient_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    lib.wf_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
    lib.wf_ctx_create.restype = C.c_int
    lib.wf_ctx_destroy.argtypes = [P]
    lib.wf_ctx_destroy.restype = None
    lib.wf_device_query.argtypes = [P, C.POINTER(wf_device_info)]
    lib.wf_device_query.restype = C.c_int
    lib.wf_strerror.argtypes = [C.c_int]
    lib.wf_strerror.restype = C.c_char_p
    lib.wf_abi_version.argtypes = []
    lib.wf_abi_version.restype = C.c_int
    lib.wf_ctx_set_timeout.argtypes = [P, C.c_double]
    lib.wf_ctx_set_timeout.restype = C.c_int
    lib.wf_align.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                             C.POINTER(wf_tile), C.POINTER(wf_result)]
    lib.wf_align.restype = C.c_int
    lib.wf_align_noalloc.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32),
                                     C.c_int64, C.POINTER(wf_tile), C.POINTER(wf_result)]
    lib.wf_align_noalloc.restype = C.c_int
    lib.wf_align_table.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                                   C.POINTER(wf_tile), C.POINTER(C.c_int32), C.POINTER(wf_result)]
    lib.wf_align_table.restype = C.c_int
    lib.wf_align_naive.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                                   C.POINTER(wf_tile), C.c_int, C.POINTER(C.c_int32), C.POINTER(wf_result)]
    lib.wf_align_naive.restype = C.c_int
    lib.wf_stencil.argtypes = [P, C.c_int, C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                               C.POINTER(wf_tile), C.POINTER(wf_result)]
    lib.wf_stencil.restype = C.c_int
    lib.wf_stencil_device.argtypes = [P, C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                      C.POINTER(wf_tile), C.POINTER(wf_result)]
    lib.wf_stencil_device.restype = C.c_int
    lib.wf_align_batch.argtypes = [P, C.POINTER(wf_job), C.c_int64, C.POINTER(wf_tile), C.POINTER(wf_result)]
    lib.wf_align_batch.restype = C.c_int
    lib.wf_plan_create.argtypes = [P, C.POINTER(wf_job), C.c_int64, C.POINTER(wf_tile), C.POINTER(P)]
    lib.wf_plan_create.restype = C.c_int
    lib.wf_plan_run.argtypes = [P, C.c_int]
    lib.wf_plan_run.restype = C.c_int
    lib.wf_plan_capture.argtypes = [P, C.c_int]
    lib.wf_plan_capture.restype = C.c_int
    lib.wf_autotune.argtypes = [P, C.POINTER(wf_job), C.c_int64, C.c_int, C.POINTER(wf_tile), C.POINTER(C.c_double)]
    lib.wf_autotune.restype = C.c_int
    lib.wf_tune_load.argtypes = [C.c_char_p]
    lib.wf_tune_load.restype = C.c_int
    lib.wf_tune_save.argtypes = [C.c_char_p]
    lib.wf_tune_save.restype = C.c_int
    lib.wf_tune_lookup_tile.argtypes = [P, C.POINTER(wf_job), C.c_int64, C.POINTER(wf_tile)]
    lib.wf_tune_lookup_tile.restype = C.c_int
    lib.wf_tune_clear.argtypes = []
    lib.wf_tune_clear.restype = C.c_int
    lib.wf_plan_results.argtypes = [P, C.POINTER(wf_result), C.c_int64]
    lib.wf_plan_results.restype = C.c_int
    lib.wf_plan_kernel_ms.argtypes = [P]
    lib.wf_plan_kernel_ms.restype = C.c_double
    lib.wf_plan_describe.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    lib.wf_variant_count.argtypes = []
    lib.wf_variant_count.restype = C.c_int
    lib.wf_variant_info.argtypes = [C.c_int] + [C.POINTER(C.c_int32)] * 4
    lib.wf_variant_info.restype = C.c_int
    lib.wf_plan_describe.restype = C.c_int
    lib.wf_plan_destroy.argtypes = [P]
    lib.wf_plan_destroy.restype = None
    _bind_server(lib)
    _lib = lib
    return lib


def _bind_server(lib: C.CDLL) -> None:
    """Bind the serving entry points."""
    from . import _server_abi
    _server_abi.bind(lib)


def strerror(status: int) -> str:
    try:
        return load().wf_strerror(status).decode()
    except ImportError:
        return f"status {status}"


def check(status: int, what: str = "") -> None:
    if status != WF_OK:
        raise WfError(status, what)
