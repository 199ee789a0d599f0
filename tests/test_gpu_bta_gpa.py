"""GPU parity of the device radix select, the bitonic top-k (BTA) and the
iterative prune+compact (GPA) paths, and the reference-style C++ driver."""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
import topk_amd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RNG = np.random.default_rng(17)


@pytest.mark.parametrize("n", [1, 1000, 1 << 20, 5_000_011])
def test_radix_select_matches_oracle(gpu, n):
    keys = RNG.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 1000:
        keys[: n // 3] = keys[0]  # a heavy duplicate
    t = torch.from_numpy(keys.view(np.int32)).cuda()
    torch.cuda.synchronize()  # the engine reads the buffer on its own stream
    with topk_amd.Engine(np.zeros((1, 10), np.float32)) as eng:
        for r in sorted({0, n // 2, n - 1, max(0, n - 64)}):
            got = eng.radix_select_u32(t.data_ptr(), n, r)
            assert got == oracle.radix_select_u32(keys, r) == np.sort(keys)[r]


def test_radix_select_float_threshold(gpu):
    """The reference use: the k-th largest non-negative score as a bit pattern
    (gpu/GPA.h:207-214): rank n - k of the float bits."""
    n, k = 2_000_000, 64
    s = RNG.random(n, dtype=np.float32)
    t = torch.from_numpy(s.view(np.int32)).cuda()
    torch.cuda.synchronize()
    with topk_amd.Engine(np.zeros((1, 10), np.float32)) as eng:
        bits = eng.radix_select_u32(t.data_ptr(), n, n - k)
    assert np.uint32(bits).view(np.float32) == np.sort(s)[n - k]


def _check(eng, fn, x, n, k, w, attr):
    ids, sc, thr = fn(k, w, attr)
    rid, rsc, rthr = oracle.fte_topk(x, n, k, w, attr)
    assert np.array_equal(ids, rid), (fn.__name__, n, k, list(attr))
    assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32))
    assert thr == rthr


@pytest.mark.parametrize("distr", [0, 1, 2])
def test_bta_and_gpa_parity(gpu, distr):
    n, d = 1_000_003, 8
    x = topk_amd.generate(n, d, distr, 60 + distr)
    with topk_amd.Engine(x) as eng:
        for k in (1, 16, 64, 100, 1024):
            for m in (2, 8):
                attr = list(range(d - m, d))
                _check(eng, eng.bta, x, n, k, np.ones(d, np.float32), attr)
                _check(eng, eng.gpa, x, n, k, np.ones(d, np.float32), attr)
        w = RNG.uniform(-1, 1, d).astype(np.float32)
        for attr in ([5, 1, 7], [0, 3, 3, 2], [7, 6, 5, 4, 3, 2, 1, 0]):
            _check(eng, eng.bta, x, n, 77, w, attr)
            _check(eng, eng.gpa, x, n, 77, w, attr)


@pytest.mark.parametrize("n", [1, 5, 4095, 4097, 70_000])
def test_bta_gpa_small(gpu, n):
    d = 3
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, n + 1)
    with topk_amd.Engine(x) as eng:
        for k in (1, 8, 64):
            _check(eng, eng.bta, x, n, k, np.ones(d, np.float32), [0, 2])
            _check(eng, eng.gpa, x, n, k, np.ones(d, np.float32), [2, 0, 1])


def test_bta_gpa_ties(gpu):
    n, d = 300_000, 4
    x = (RNG.integers(0, 3, size=(d, n)) / 2.0).astype(np.float32)
    with topk_amd.Engine(x) as eng:
        for k in (10, 500):
            _check(eng, eng.bta, x, n, k, np.ones(d, np.float32), [0, 1, 2, 3])
            _check(eng, eng.gpa, x, n, k, np.array([1, -0.5, 2, 0.25], np.float32), [3, 1, 0])


def test_gpa_on_reordered_store(gpu):
    """The reference's own GPA setting: reorder_attr, then all attributes
    with unit weights."""
    n, d = 500_000, 6
    x = topk_amd.generate(n, d, topk_amd.CORRELATED, 9)
    with topk_amd.Engine(x) as eng:
        eng.reorder_attr()
        y = oracle.reorder_attr(x, n)
        _check(eng, eng.gpa, y, n, 32, np.ones(d, np.float32), list(range(d)))
        st = eng.stats()
        assert st["candidates"] >= 32


def test_cpp_driver_check_and_sweep(gpu):
    exe = os.path.join(ROOT, "topk_amd", "bin", "bench_topk")
    r = subprocess.run([exe, "-n", "300000", "-d", "6", "-s", "16", "-e", "64", "--check"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAILED" not in r.stdout and r.stdout.count("PASSED") == 3 * 5
    r = subprocess.run([exe, "-n", "200000", "-d", "4", "-a", "pta", "-i", "2", "--no-normalize"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    for key in ("tt_init:", "tt_procesing:", "tuple_count:", "threshold:"):
        assert key in r.stdout
    r = subprocess.run([exe, "-n", "200000", "-d", "8", "-a", "mq", "-w", "64", "-i", "1"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "tuples_per_second:" in r.stdout
