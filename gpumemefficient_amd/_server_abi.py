"""ctypes declarations of the serving entry points (wf_server_*)."""
from __future__ import annotations

import ctypes as C

SERVER_EXPORTS = ["wf_server_create", "wf_server_submit", "wf_server_poll", "wf_server_wait_idle",
                  "wf_server_stats", "wf_server_replay", "wf_server_destroy"]


class wf_server_opts(C.Structure):
    _fields_ = [("streams", C.c_int32), ("max_batch", C.c_int32), ("hbm_budget", C.c_int64),
                ("queue_capacity", C.c_int32), ("per_request", C.c_int32), ("resident", C.c_int32)]


class wf_completion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("last", C.c_int32), ("max_score", C.c_int32), ("status", C.c_int32),
                ("reserved", C.c_int32), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class wf_server_stats_t(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("completed", C.c_uint64), ("cells", C.c_uint64),
                ("hbm_budget", C.c_int64), ("hbm_used", C.c_int64), ("pending", C.c_int64)]


def bind(lib: C.CDLL) -> None:
    """Declare argtypes for the wf_server_* functions of `lib`."""
    from ._lib import wf_job
    P = C.c_void_p
    lib.wf_server_create.argtypes = [C.c_int, C.POINTER(wf_server_opts), C.POINTER(P)]
    lib.wf_server_create.restype = C.c_int
    lib.wf_server_submit.argtypes = [P, C.POINTER(wf_job), C.POINTER(C.c_uint64)]
    lib.wf_server_submit.restype = C.c_int
    lib.wf_server_poll.argtypes = [P, C.POINTER(wf_completion), C.c_int, C.POINTER(C.c_int)]
    lib.wf_server_poll.restype = C.c_int
    lib.wf_server_wait_idle.argtypes = [P, C.c_double]
    lib.wf_server_wait_idle.restype = C.c_int
    lib.wf_server_stats.argtypes = [P, C.POINTER(wf_server_stats_t)]
    lib.wf_server_stats.restype = C.c_int
    lib.wf_server_replay.argtypes = [P, C.POINTER(wf_job), C.POINTER(C.c_double), C.c_int64,
                                     C.POINTER(wf_completion)]
    lib.wf_server_replay.restype = C.c_int
    lib.wf_server_destroy.argtypes = [P]
    lib.wf_server_destroy.restype = None
