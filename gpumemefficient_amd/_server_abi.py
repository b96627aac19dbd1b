"""ctypes declarations of the serving entry points (wf_server_*)."""
from __future__ import annotations

import ctypes as C

SERVER_EXPORTS: list[str] = []


def bind(lib: C.CDLL) -> None:
    """Declare argtypes for the wf_server_* functions present in `lib`."""
    return None
