import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def ctx():
    """One libwfdp context for the whole GPU session.  A missing library is an
    error (the product path never falls back); only a host without any HIP
    device skips."""
    from gpumemefficient_amd import Context, WfError
    from gpumemefficient_amd import _lib
    _lib.load()
    try:
        c = Context(0, timeout_s=20.0)
    except WfError as e:
        if e.status == _lib.WF_ERR_DEVICE:
            pytest.skip("no HIP device on this host")
        raise
    yield c
    c.close()
