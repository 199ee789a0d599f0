"""CPU-side checks of the C ABI library: it loads, exports every function
include/topk.h declares, reports errors as status codes, and its host-side
generator is deterministic and has the documented distributions."""
import ctypes
import os
import re

import numpy as np
import pytest

import topk_amd
from topk_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "topk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(topk_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_expected_entry_points():
    names = _declared_functions()
    for n in ("topk_ctx_create", "topk_fte", "topk_pta", "topk_batch", "topk_ctx_destroy", "topk_init"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared_functions()) == set(_lib.SIGNATURES)


def test_abi_version_and_strerror():
    L = topk_amd.lib()
    assert L.topk_abi_version() == 1
    for s in range(0, 9):
        assert L.topk_strerror(s)


def test_invalid_arguments_return_status():
    L = topk_amd.lib()
    h = ctypes.c_void_p()
    data = np.zeros((2, 4), np.float32)
    assert L.topk_ctx_create(ctypes.byref(h), None, 4, 2, 0) == _lib.TOPK_ERR_INVALID
    assert L.topk_ctx_create(ctypes.byref(h), data.ctypes.data_as(ctypes.c_void_p), 4, 0, 0) == _lib.TOPK_ERR_INVALID
    assert L.topk_ctx_create(ctypes.byref(h), data.ctypes.data_as(ctypes.c_void_p), 4, 33, 0) == _lib.TOPK_ERR_INVALID
    assert L.topk_fte(None, 4, 1, None, None, None, None, None) == _lib.TOPK_ERR_INVALID
    assert L.topk_ctx_destroy(None) == _lib.TOPK_ERR_INVALID


def test_no_device_is_reported(monkeypatch):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine(np.ones((2, 8), np.float32))
    assert e.value.status == _lib.TOPK_ERR_NO_DEVICE


def test_generator_deterministic_and_range_consistent():
    n = 50_000
    for dist in range(4):
        a = topk_amd.generate(n, 6, dist, 7, threads=3)
        b = topk_amd.generate(n, 6, dist, 7, threads=1)
        assert np.array_equal(a, b)
        c = topk_amd.generate(n, 6, dist, 7, i0=12_345, count=777)
        assert np.array_equal(c, a[:, 12_345:12_345 + 777])
        assert np.all(a >= 0) and np.all(a <= 1)
    assert not np.array_equal(topk_amd.generate(1000, 4, 1, 1), topk_amd.generate(1000, 4, 1, 2))


def test_generator_distributions():
    n, d = 200_000, 8
    ind = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0)
    assert abs(ind.mean() - 0.5) < 0.005 and ind.max() < 1.0
    assert abs(np.corrcoef(ind[0], ind[1])[0, 1]) < 0.02
    cor = topk_amd.generate(n, d, topk_amd.CORRELATED, 0)
    assert np.corrcoef(cor[0], cor[1])[0, 1] > 0.9
    assert (cor == 1.0).mean() > 0.005  # clipping creates exact ties at 1.0
    anti = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 0)
    s = anti.astype(np.float64).sum(axis=0)
    assert abs(s.mean() - d / 2) < 0.01 and s.std() < 0.1
    assert np.corrcoef(anti[0], anti[1])[0, 1] < -0.05
    mix = topk_amd.generate(n, d, topk_amd.MIXED, 0)
    third = n // 3
    assert np.array_equal(mix[:, :third], topk_amd.generate(n, d, topk_amd.CORRELATED, 0)[:, :third])
    assert np.array_equal(mix[:, 2 * third:], anti[:, 2 * third:])
