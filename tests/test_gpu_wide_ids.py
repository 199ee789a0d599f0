"""Stores of 2^32 - 1 objects or more: 64-bit object ids (the reference's
result tuples carry uint64_t tids, cpu/AA.h:29-35).  The FTE path runs per
id shard of 2^31 objects and merges the shards' top-k lists on the device
by score and global id; results equal the oracle's 128-bit-key
restatement, including ids above 2^32 and ties that straddle the shard and
2^32 boundaries."""
import numpy as np
import pytest

import oracle
import topk_amd
from topk_amd import _lib

pytestmark = pytest.mark.gpu

N = (1 << 32) + 1000
B32 = 1 << 32


@pytest.fixture(scope="module")
def planted(gpu):
    """Zeros, a window of quantised values around 2^32 (ties inside it) and
    planted values: the maximum above 2^32, ties at 0.75 in three shards and
    at 0.5 on both sides of 2^32 (0xFFFFFFFF included)."""
    x = np.zeros((1, N), np.float32)
    rng = np.random.default_rng(77)
    lo = B32 - 100
    x[0, lo:] = (rng.integers(1, 400, N - lo) / 1024.0).astype(np.float32)
    for i, v in {B32 + 999: 1.0, 3: 0.75, B32 + 5: 0.75, (1 << 31) + 7: 0.75, (1 << 31) - 1: 0.75,
                 B32 - 2: 0.5, B32 - 1: 0.5, B32: 0.5, B32 + 1: 0.5}.items():
        x[0, i] = v
    eng = topk_amd.Engine(x)
    yield x, eng
    eng.close()


@pytest.mark.parametrize("k", [1, 16, 64, 1024])
def test_wide_fte_matches_oracle(planted, k):
    x, eng = planted
    w = np.ones(1, np.float32)
    gi, gs, gt = eng.fte(k, w, [0])
    ri, rs, rt = oracle.fte_topk_wide(x, N, k, w, [0], threads=16)
    assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rt
    assert gi[0] == B32 + 999 and gs[0] == 1.0
    if k >= 16:
        # the 0.75 ties rank by id across shards 0, 1 and 2; the 0.5 ties
        # straddle 2^32 (0xFFFFFFFF is an ordinary id here)
        assert gi[1:5].tolist() == [3, (1 << 31) - 1, (1 << 31) + 7, B32 + 5]
        assert gi[5:9].tolist() == [B32 - 2, B32 - 1, B32, B32 + 1]
    assert eng.stats()["tuple_count"] == N


def test_wide_fte_negative_weight_ties_at_zero(planted):
    """Every zero ties: the lowest ids win, the planted values rank last."""
    x, eng = planted
    w = np.full(1, -1.0, np.float32)
    gi, gs, gt = eng.fte(64, w, [0])
    ri, rs, rt = oracle.fte_topk_wide(x, N, 64, w, [0], threads=16)
    assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rt
    assert gi.tolist() == [i for i in range(65) if i != 3]


def test_wide_batch_runs_every_query_through_fte(planted):
    x, eng = planted
    qs = [(16, np.ones(1, np.float32), np.array([0], np.uint8)),
          (100, np.full(1, -0.5, np.float32), np.array([0], np.uint8))]
    for (gi, gs, gt), q in zip(eng.batch(qs), qs):
        ri, rs, rt = eng.fte(*q)
        assert np.array_equal(gi, ri) and np.array_equal(gs, rs) and gt == rt


def test_wide_paths_with_32bit_ids_refuse(planted):
    _, eng = planted
    for call in (lambda: eng.init(topk_amd.TOPK_INIT_PTA), lambda: eng.bta(8, np.ones(1, np.float32), [0])):
        with pytest.raises(topk_amd.TopkError) as e:
            call()
        assert e.value.status == _lib.TOPK_ERR_UNSUPPORTED


def test_wide_generated_async_and_sync_match_oracle(gpu):
    """A generated store: queries through both pipeline lanes (per-shard
    runs and the merge on the device) equal the synchronous path and the
    oracle on the host generator's identical values."""
    import torch
    with topk_amd.Engine.generated(N, 1, topk_amd.INDEPENDENT, 11) as eng:
        w = np.ones(1, np.float32)
        sync = {k: eng.fte(k, w, [0]) for k in (64, 256)}
        outs = []
        for k in (64, 256, 64):
            ids = torch.empty(k, dtype=torch.int64, device="cuda")
            sc = torch.empty(k, dtype=torch.float32, device="cuda")
            eng.fte_async(k, w, [0], ids.data_ptr(), sc.data_ptr())
            outs.append((k, ids, sc))
        eng.sync()
        for k, ids, sc in outs:
            gi, gs, _ = sync[k]
            assert np.array_equal(ids.cpu().numpy().view(np.uint64), gi)
            assert np.array_equal(sc.cpu().numpy(), gs)
    x = topk_amd.generate(N, 1, topk_amd.INDEPENDENT, 11, threads=16)
    ri, rs, rt = oracle.fte_topk_wide(x, N, 64, w, [0], threads=16)
    gi, gs, gt = sync[64]
    assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rt


def test_large_store_layout_caps_partitions_for_batches(gpu):
    """2^30 + 12345 objects at the default 4096-object partitions would need
    262,147 partitions: the layout stops at 2^18 (larger partitions) and
    batches, PTA and FTE agree with the oracle."""
    n, d = (1 << 30) + 12345, 2
    rng = np.random.default_rng(5)
    qs = [(k, rng.uniform(0.1, 1.0, d).astype(np.float32), np.array(a, np.uint8))
          for k, a in ((64, [0, 1]), (16, [1]), (256, [1, 0]))]
    with topk_amd.Engine.generated(n, d, topk_amd.ANTICORRELATED, 9) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        pid, poff, psz, pmax, pmin = eng.pta_layout()
        assert psz.size == 1 << 18 and int(psz.sum()) == n
        got = eng.batch(qs)
        pta = [eng.pta(*q) for q in qs]
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 9, threads=16)
    for q, (gi, gs, gt), (pi, ps, pt) in zip(qs, got, pta):
        ri, rs, rt = oracle.tpac_threads(x, n, q[0], q[1], q[2], 16)
        assert np.array_equal(gi, ri) and np.array_equal(gs.view(np.uint32), rs.view(np.uint32)) and gt == rt
        assert np.array_equal(pi, ri) and np.array_equal(ps, rs) and pt == rt
