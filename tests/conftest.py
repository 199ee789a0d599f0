import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs on the GPU box)")


def gpu_available():
    """Asks the library itself (not torch): a context on a one-object store
    either builds or reports TOPK_ERR_NO_DEVICE.  A missing library raises."""
    import ctypes

    import numpy as np

    from topk_amd import _lib
    L = _lib.lib()
    x = np.zeros(1, np.float32)
    h = ctypes.c_void_p()
    rc = L.topk_ctx_create(ctypes.byref(h), x.ctypes.data, 1, 1, 0)
    if rc == _lib.TOPK_OK:
        L.topk_ctx_destroy(h)
        return True
    if rc == _lib.TOPK_ERR_NO_DEVICE:
        return False
    raise RuntimeError(f"topk_ctx_create on a one-object store failed with status {rc}")


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU (topk_ctx_create reports TOPK_ERR_NO_DEVICE)")
    return True
