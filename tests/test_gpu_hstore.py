"""Stores in pinned host memory streamed through HBM per query (the
reference's GPAm role, gpu/GPAm.h:30-89): every answer equals the in-HBM
FTE path and the oracle bit for bit, across chunk boundaries, ragged last
chunks, ties that straddle chunks, and batches that stream each chunk once."""
import numpy as np
import pytest

import oracle
import topk_amd
from topk_amd import _lib
from topk_amd.engine import _ptr

pytestmark = pytest.mark.gpu


def _same(a, b):
    (ai, asc, at), (bi, bsc, bt) = a, b
    return (np.array_equal(ai, bi) and np.array_equal(asc.view(np.uint32), bsc.view(np.uint32))
            and np.float32(at).tobytes() == np.float32(bt).tobytes())


def _host(x, chunk):
    d, n = x.shape
    hs = topk_amd.HostStore(n, d, chunk)
    hs.data[:] = x
    return hs


def test_streamed_equals_resident(gpu):
    n, d = 1_000_003, 8
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 21)
    with topk_amd.Engine(x) as e, _host(x, 1 << 18) as hs:
        rng = np.random.default_rng(0)
        for k in (1, 64, 1024):
            for at in (list(range(d)), [2, 5], [7, 0, 3]):
                w = rng.uniform(0.1, 1.0, d).astype(np.float32)
                assert _same(hs.fte(k, w, at), e.fte(k, w, at)), (k, at)
        rep = hs.report()
        assert rep["chunks"] == 4 and rep["bytes"] == 4 * n * 3
    w = np.linspace(1.0, 0.3, d).astype(np.float32)
    with _host(x, 1 << 18) as hs:
        ri, rs, rt = oracle.fte_topk(x, n, 100, w, list(range(d)))
        assert _same(hs.fte(100, w, list(range(d))), (ri, rs, rt))


def test_ties_across_chunks(gpu):
    n, d = 40_000, 3
    rng = np.random.default_rng(3)
    x = (rng.integers(0, 3, (d, n)) * 0.5).astype(np.float32)   # few distinct scores: ties everywhere
    with topk_amd.Engine(x) as e, _host(x, 8192) as hs:
        for k in (10, 500, 1024):
            w = np.ones(d, np.float32)
            got = hs.fte(k, w, [0, 1, 2])
            assert _same(got, e.fte(k, w, [0, 1, 2]))
            assert np.all(np.diff(got[0][got[1] == got[1][0]].astype(np.int64)) > 0)  # equal scores: lower id first


def test_small_stores_and_short_tails(gpu):
    x = topk_amd.generate(700, 4, topk_amd.CORRELATED, 2)
    with topk_amd.Engine(x) as e, _host(x, 8192) as hs:      # k > n: padded like topk_fte
        w = np.ones(4, np.float32)
        assert _same(hs.fte(1024, w, [0, 1, 2, 3]), e.fte(1024, w, [0, 1, 2, 3]))
    x = topk_amd.generate(8192 + 100, 4, topk_amd.ANTICORRELATED, 2)
    with topk_amd.Engine(x) as e, _host(x, 8192) as hs:      # last chunk holds fewer than k objects
        w = np.array([0.3, 0.9, 0.1, 0.5], np.float32)
        assert _same(hs.fte(1024, w, [3, 1]), e.fte(1024, w, [3, 1]))


def test_batch_streams_each_chunk_once(gpu):
    n, d = 600_000, 12
    x = topk_amd.generate(n, d, topk_amd.MIXED, 8)
    rng = np.random.default_rng(8)
    queries = []
    for _ in range(40):
        m = int(rng.integers(1, d + 1))
        at = [int(a) for a in rng.permutation(d)[:m]]
        queries.append((int(rng.choice([1, 16, 64, 256, 1024])), rng.uniform(-1, 1, d).astype(np.float32), at))
    with topk_amd.Engine(x) as e, _host(x, 1 << 17) as hs:
        got = hs.batch(queries)
        rep = hs.report()
        used = sorted({a for _, _, at in queries for a in at})
        assert rep["queries"] == 40 and rep["bytes"] == 4 * n * len(used)
        for q, g in zip(queries, got):
            assert _same(g, e.fte(*q))


def test_errors(gpu):
    x = topk_amd.generate(20_000, 4, topk_amd.INDEPENDENT, 1)
    with _host(x, 8192) as hs:
        hs.data[2, 12_345] = np.nan
        w = np.ones(4, np.float32)
        assert hs.fte(8, w, [0, 1])[0].size == 8             # NaN in a column no query streams
        with pytest.raises(topk_amd.TopkError) as e:
            hs.fte(8, w, [2])
        assert e.value.status == _lib.TOPK_ERR_NONFINITE
        with pytest.raises(topk_amd.TopkError) as e:
            hs.fte(8, w, [4])
        assert e.value.status == _lib.TOPK_ERR_INVALID
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.HostStore(20_000, 4, 100)
    assert e.value.status == _lib.TOPK_ERR_INVALID


def test_full_size_host_store(gpu):
    """N = 2^27, d = 8 (4 GiB pinned host store), chunks of 2^25 objects."""
    n, d = 1 << 27, 8
    with topk_amd.Engine.generated(n, d, topk_amd.INDEPENDENT, 0) as e, topk_amd.HostStore(n, d, 1 << 25) as hs:
        hs.data[:] = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0, threads=16)
        w = np.ones(d, np.float32)
        assert _same(hs.fte(64, w, list(range(d))), e.fte(64, w, list(range(d))))
        rep = hs.report()
        assert rep["chunks"] == 4 and rep["bytes"] == 4 * n * d


def test_weights_read_only_at_queried_attributes(gpu):
    # weights are indexed by attribute id; an array sized to max(attr) + 1 is
    # enough, as for topk_fte (entries past it are never read)
    n, d = 50_000, 8
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 4)
    at = np.array([2, 0], np.uint8)
    w_short = np.array([0.5, 9.0, 0.25], np.float32)   # attribute 1 is not queried
    k = 20
    ids = np.empty(k, np.uint64)
    sc = np.empty(k, np.float32)
    thr = np.zeros(1, np.float32)
    with _host(x, 1 << 14) as hs:
        _lib.check(_lib.lib().topk_hstore_fte(hs._h, k, 2, _ptr(w_short),
                                                  _ptr(at), _ptr(ids),
                                                  _ptr(sc), _ptr(thr)),
                       "topk_hstore_fte")
        w_full = np.zeros(d, np.float32)
        w_full[[0, 2]] = [0.5, 0.25]
        ri, rs, rt = oracle.fte_topk(x, n, k, w_full, [2, 0])
        assert _same((ids, sc, float(thr[0])), (ri, rs, rt))
        bad = np.array([9], np.uint8)
        with pytest.raises(topk_amd.TopkError):
            _lib.check(_lib.lib().topk_hstore_fte(hs._h, k, 1, _ptr(w_short),
                                                      _ptr(bad), _ptr(ids),
                                                      _ptr(sc), _ptr(thr)),
                           "topk_hstore_fte")
