"""Python handle on the continuous-batching server (wf_server_* in the C
ABI) and the serving benchmark's request loop."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable

import numpy as np

from . import _lib
from ._lib import check
from ._server_abi import wf_completion, wf_server_opts, wf_server_stats_t
from .api import Job, _JobArray


@dataclass
class Completion:
    ticket: int
    last: int
    max_score: int
    status: int
    t_submit: float
    t_start: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


def _comp(c: wf_completion) -> Completion:
    return Completion(int(c.ticket), int(c.last), int(c.max_score), int(c.status), float(c.t_submit),
                      float(c.t_start), float(c.t_done))


class Server:
    def __init__(self, device: int = 0, streams: int = 4, max_batch: int = 512, hbm_budget: int = 0,
                 queue_capacity: int = 0, per_request: bool = False, resident: bool = False):
        """per_request: one launch per request (baseline); resident: one
        persistent launch for the server's lifetime that takes requests in
        as they arrive (device-resident task queue); else launches of the
        admitted requests on `streams` slots."""
        self._lib = _lib.load()
        self._h = C.c_void_p()
        opts = wf_server_opts(streams, max_batch, hbm_budget, queue_capacity, 1 if per_request else 0,
                              1 if resident else 0)
        check(self._lib.wf_server_create(device, C.byref(opts), C.byref(self._h)), "wf_server_create")
        self._keep = []

    def submit(self, job: Job) -> int:
        ja = _JobArray([job])
        t = C.c_uint64()
        check(self._lib.wf_server_submit(self._h, ja.c, C.byref(t)), "wf_server_submit")
        return int(t.value)

    def poll(self, max_n: int = 4096) -> list[Completion]:
        """Completions since the last poll.  After a device error the
        completions taken off the queue ride on the raised WfError
        (``.completions``) instead of being lost."""
        buf = (wf_completion * max_n)()
        n = C.c_int()
        st = self._lib.wf_server_poll(self._h, buf, max_n, C.byref(n))
        comps = [_comp(buf[i]) for i in range(n.value)]
        if st != _lib.WF_OK:
            err = _lib.WfError(st, "wf_server_poll")
            err.completions = comps
            raise err
        return comps

    def wait_idle(self, timeout_s: float = 60.0) -> None:
        check(self._lib.wf_server_wait_idle(self._h, timeout_s), "wf_server_wait_idle")

    def stats(self) -> dict:
        st = wf_server_stats_t()
        check(self._lib.wf_server_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in wf_server_stats_t._fields_}

    def replay(self, jobs: Iterable[Job], arrival_s) -> list[Completion]:
        """Submit jobs[i] at arrival_s[i] seconds (real time) and wait for
        all; returns completions in job order (t_submit = scheduled arrival)."""
        ja = _JobArray(jobs)
        arr = np.ascontiguousarray(np.asarray(arrival_s, dtype=np.float64))
        out = (wf_completion * max(1, ja.n))()
        check(self._lib.wf_server_replay(self._h, ja.c, arr.ctypes.data_as(C.POINTER(C.c_double)), ja.n, out),
              "wf_server_replay")
        return [_comp(out[i]) for i in range(ja.n)]

    def close(self) -> None:
        if self._h:
            self._lib.wf_server_destroy(self._h)
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


def percentile(values, q: float) -> float:
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    return float(np.percentile(v, q))
