"""Parity at the full workload sizes of the partitioned, batched and learned
paths (the headline FTE shape, N = 2^26 / d = 8, is covered in
test_gpu_fte.py).

Each test compares against the oracle's threaded scans
(cpu/TPAc.h:256-358 findTopKthreads, cpu/TPAc.h:209-254 findTopKsimdMQ) on
the identical buffer, for as many queries as the oracle finishes in seconds,
and checks the rest through a size-independent property: every answer of
the pruned / batched / learned path equals the exact FTE answer of the same
query, bit for bit."""
import numpy as np
import pytest

import oracle
import topk_amd

from test_gpu_batch import make_queries, oracle_batch

pytestmark = pytest.mark.gpu

HOST_THREADS = 16
KS = [16, 32, 64, 128, 256]


def _same(a, b):
    (ai, asc, at), (bi, bsc, bt) = a, b
    return np.array_equal(ai, bi) and np.array_equal(asc.view(np.uint32), bsc.view(np.uint32)) and at == bt


def test_pta_full_size_anticorrelated(gpu):
    """N = 2^28, d = 8 anticorrelated (seed 3), 4096-object partitions,
    weights U(0.1, 1) on all 8 attributes (query seed 100), k = 128."""
    n, d, k = 1 << 28, 8, 128
    rng = np.random.default_rng(100)
    ws = [rng.uniform(0.1, 1.0, d).astype(np.float32) for _ in range(3)]
    attr = list(range(d))
    with topk_amd.Engine.generated(n, d, topk_amd.ANTICORRELATED, 3) as eng:
        eng.set_partition_size(4096)
        eng.init(topk_amd.TOPK_INIT_PTA)
        got, scored = [], []
        for w in ws:
            got.append(eng.pta(k, w, attr))
            st = eng.stats()
            scored.append((st["partitions_scored"], st["partitions_total"]))
            assert _same(got[-1], eng.fte(k, w, attr))
    assert all(s <= t and t >= (n // 4096) // 2 for s, t in scored)
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 3, threads=HOST_THREADS)
    for w, (ids, sc, thr) in zip(ws, got):
        rid, rsc, rthr = oracle.tpac_threads(x, n, k, w, attr, HOST_THREADS)
        assert np.array_equal(ids, rid) and np.array_equal(sc.view(np.uint32), rsc.view(np.uint32)) and thr == rthr


def test_batch_full_size_d16_mixed(gpu):
    """N = 2^25, d = 16 mixed (seed 4), one launch of 1024 queries on random
    subsets of 2-16 attributes, weights U(0.1, 1), k from {16..256} (query
    seed 200), after the launch-shape autotuner has picked its shape."""
    n, d = 1 << 25, 16
    rng = np.random.default_rng(200)
    qs = make_queries(rng, 1024, d, KS)
    with topk_amd.Engine.generated(n, d, topk_amd.MIXED, 4) as eng:
        tune = eng.batch_autotune(qs)
        assert tune["ms_best"] > 0
        got = eng.batch(qs)
        st = eng.stats()
        assert st["queries"] == 1024 and st["fallback"] <= 0.01 * len(qs)
        for i, q in enumerate(qs):
            assert _same(got[i], eng.fte(*q)), i
    x = topk_amd.generate(n, d, topk_amd.MIXED, 4, threads=HOST_THREADS)
    sub = list(range(0, 1024, 32))
    ref = oracle_batch(x, n, [qs[i] for i in sub], threads=HOST_THREADS)
    for i, (ri, rs) in zip(sub, ref):
        gi, gs, gt = got[i]
        assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rs[-1], i


def test_learned_full_size_d16_mixed(gpu):
    """N = 2^26, d = 16 mixed (seed 5): the predictor is trained on 100k
    queries (query seed 300) labelled by the GPU's exact path and evaluated on
    4096 held-out queries (seed 301) in four launches of 1024."""
    n, d = 1 << 26, 16
    train = make_queries(np.random.default_rng(300), 100_000, d, KS)
    held = make_queries(np.random.default_rng(301), 4096, d, KS)
    got, fallback, fte = [], 0, 0
    with topk_amd.Engine.generated(n, d, topk_amd.MIXED, 5) as eng:
        rep = eng.learned_train(train)
        assert rep["n_train"] + rep["n_calib"] == len(train)
        assert rep["margin"] > 0 and rep["calib_fallback"] <= 0.01
        eng.set_tuning("batch_mode", 2)
        for b in range(0, len(held), 1024):
            got += eng.batch(held[b:b + 1024])
            st = eng.stats()
            fallback += st["fallback"]
            fte += st["fte_fallback"]
        eng.set_tuning("batch_mode", 0)
        assert fallback <= 0.02 * len(held) and fte <= 0.005 * len(held)
        mism = sum(not _same(g, eng.fte(*q)) for g, q in zip(got, held))
        assert mism == 0
    x = topk_amd.generate(n, d, topk_amd.MIXED, 5, threads=HOST_THREADS)
    sub = list(range(0, len(held), 128))
    ref = oracle_batch(x, n, [held[i] for i in sub], threads=HOST_THREADS)
    for i, (ri, rs) in zip(sub, ref):
        gi, gs, gt = got[i]
        assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rs[-1], i
