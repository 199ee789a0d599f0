"""The query-major fused batch pass (topk_set_tuning "batch_fused" 2: one
workgroup per query bounds groups, partitions and chunks and scores the
chunks that reach its threshold) against the oracle's multi-query
restatement (cpu/TPAc.h:209-254), over the same cases as the work-item
path: distributions, widths, term orders with repeats and negative
weights, tiny stores, tie overflows, graph replays, pipelined batches and
learned thresholds with forced fallbacks."""
import numpy as np
import pytest

import topk_amd

from test_gpu_batch import check_batch, make_queries, oracle_batch

pytestmark = pytest.mark.gpu


def fused(eng):
    eng.set_tuning("batch_fused", 2)
    return eng


@pytest.mark.parametrize("distr", [0, 1, 2, 3])
def test_fused_parity_d8(gpu, distr):
    n, d = 1_000_003, 8
    rng = np.random.default_rng(300 + distr)
    x = topk_amd.generate(n, d, distr, 70 + distr)
    with topk_amd.Engine(x) as eng:
        st = check_batch(fused(eng), x, n, make_queries(rng, 256, d, [1, 16, 32, 64, 128, 256, 1024]))
        assert st["queries"] == 256 and st["fallback"] == 0
        assert 0 < st["partitions_scored"] < st["partitions_total"]
        assert st["items"] == 0  # no work list


def test_fused_parity_d16_mixed(gpu):
    n, d = 1 << 20, 16
    rng = np.random.default_rng(6)
    x = topk_amd.generate(n, d, topk_amd.MIXED, 13)
    with topk_amd.Engine(x) as eng:
        st = check_batch(fused(eng), x, n, make_queries(rng, 1024, d, [16, 32, 64, 128, 256]))
        assert st["fallback"] <= 0.01 * 1024


@pytest.mark.parametrize("order", ["shuffled", "repeats"])
def test_fused_term_order_and_repeats(gpu, order):
    n, d = 300_000, 12
    rng = np.random.default_rng(19)
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 22)
    with topk_amd.Engine(x) as eng:
        qs = make_queries(rng, 128, d, [8, 64, 300], order=order, mmin=1, mmax=20 if order == "repeats" else d,
                          wlo=-1.0, whi=1.0)
        check_batch(fused(eng), x, n, qs)


def test_fused_wide_store(gpu):
    n, d = 200_000, 32
    rng = np.random.default_rng(23)
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 24)
    with topk_amd.Engine(x) as eng:
        check_batch(fused(eng), x, n, make_queries(rng, 64, d, [10, 100], order="shuffled"))


@pytest.mark.parametrize("n", [1, 7, 1000, 5000])
def test_fused_small_stores(gpu, n):
    d = 4
    rng = np.random.default_rng(n + 1)
    x = topk_amd.generate(n, d, topk_amd.CORRELATED, n + 1)
    with topk_amd.Engine(x) as eng:
        qs = make_queries(rng, 40, d, [1, 5, 64, 1024], mmin=1)
        got = fused(eng).batch(qs)
        for (gi, gs, _), (ri, rs) in zip(got, oracle_batch(x, n, qs)):
            assert np.array_equal(gi, ri) and np.array_equal(gs, rs)


def test_fused_tie_overflow_reruns_with_key_threshold(gpu):
    n, d = 200_000, 2
    x = np.zeros((d, n), np.float32)
    x[:, ::2] = 1.0
    with topk_amd.Engine(x) as eng:
        qs = [(100, np.ones(d, np.float32), np.array([0, 1], np.uint8)),
              (10, np.array([1.0, -1.0], np.float32), np.array([1, 0], np.uint8)),
              (1024, np.ones(d, np.float32), np.array([1], np.uint8))]
        st = check_batch(fused(eng), x, n, qs)
        assert st["fallback"] == 0


def test_fused_graph_replays_and_pipeline(gpu):
    """Graph replays with eager work between them, and two batches in
    flight, equal eager runs of the work-item path."""
    n, d = 700_000, 8
    rng = np.random.default_rng(45)
    x = topk_amd.generate(n, d, topk_amd.MIXED, 45)
    with topk_amd.Engine(x) as eng:
        pbs = [eng.prepare_batch(make_queries(rng, 300, d, [16, 64, 128])) for _ in range(4)]
        eng.set_tuning("batch_fused", 1)
        ref = [eng.batch(pb.queries) for pb in pbs]
        fused(eng)
        for _ in range(2):
            for pb, r in zip(pbs[:2], ref[:2]):
                got = eng.batch(pb.queries)
                eng.fte(8, np.ones(d, np.float32), [0, 1])
                for (gi, gs, gt), (ri, rs, rt) in zip(got, r):
                    assert np.array_equal(gi, ri) and np.array_equal(gs, rs) and gt == rt
        tickets = [eng.batch_submit(pb) for pb in pbs]
        eng.batch_wait(tickets[-1])
        for pb, r in zip(pbs, ref):
            for i, (ri, rs, rt) in enumerate(r):
                gi, gs, gt = pb.result(i, n)
                assert np.array_equal(gi, ri) and np.array_equal(gs, rs) and gt == rt


def test_fused_learned_thresholds_exact(gpu):
    """Learned thresholds through the fused pass: exact answers, and a
    predictor forced above every k-th score reruns every query exactly."""
    n, d = 1 << 20, 16
    rng = np.random.default_rng(99)
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 98)
    with topk_amd.Engine(x) as eng:
        fused(eng)
        eng.learned_train(make_queries(rng, 4000, d, [16, 32, 64, 128, 256]), epochs=30)
        test = make_queries(rng, 512, d, [16, 32, 64, 128, 256])
        eng.set_tuning("batch_mode", 2)
        got = eng.batch(test)
        ref = oracle_batch(x, n, test)
        for i, ((gi, gs, gt), (ri, rs)) in enumerate(zip(got, ref)):
            assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)), i
            assert gt == rs[-1]
        params, margin = eng.learned_get()
        eng.learned_set(params, -1.0)
        sub = test[:40]
        got = eng.batch(sub)
        st = eng.stats()
        eng.learned_set(params, margin)
        assert st["fallback"] == len(sub)
        for (gi, gs, _), (ri, rs) in zip(got, ref[:40]):
            assert np.array_equal(gi, ri) and np.array_equal(gs, rs)
