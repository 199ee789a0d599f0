"""GPU parity of the PTA path (topk_init(TOPK_INIT_PTA) + topk_pta), the
partition layout invariants, store normalisation and reorder_attr."""
import numpy as np
import pytest

import oracle
import topk_amd

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(7)


def _check_pta(eng, x, n, k, w, attr, layout=None):
    ids, sc, thr = eng.pta(k, w, attr)
    rid, rsc, rthr = oracle.fte_topk(x, n, k, w, attr)
    assert np.array_equal(ids, rid), (n, k, list(attr))
    assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32))
    assert thr == rthr
    if layout is not None:
        pdata, pid, poff, psz, pmax, pmin = layout
        lid, lsc, _, _ = oracle.pta_topk(pdata, pid, poff, psz, pmax, pmin, k, w, attr)
        assert np.array_equal(lid, ids) and np.array_equal(lsc, sc)
    return eng.stats()


def _layout(eng, x):
    pid, poff, psz, pmax, pmin = eng.pta_layout()
    pdata = np.zeros((x.shape[0], pid.size), np.float32)
    valid = pid != 0xFFFFFFFF
    pdata[:, valid] = x[:, pid[valid]]
    return pdata, pid, poff, psz, pmax, pmin


@pytest.mark.parametrize("distr", [0, 1, 2, 3])
def test_pta_layout_invariants(gpu, distr):
    n, d = 300_001, 8
    x = topk_amd.generate(n, d, distr, 30 + distr)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        pid, poff, psz, pmax, pmin = eng.pta_layout()
        valid = pid != 0xFFFFFFFF
        assert np.array_equal(np.sort(pid[valid]), np.arange(n, dtype=np.uint32))  # a permutation
        assert np.all(poff % 256 == 0) and np.all(np.diff(poff) >= psz)
        for p in range(psz.size):
            seg = pid[poff[p]:poff[p] + psz[p]]
            assert np.all(seg != 0xFFFFFFFF) and np.all(pid[poff[p] + psz[p]:poff[p + 1]] == 0xFFFFFFFF)
            # inside a partition: attribute sums (float32, attribute order)
            # non-increasing, equal sums by ascending id
            sm = np.zeros(seg.size, np.float32)
            for m in range(d):
                sm = sm + x[m, seg]
            ds, di = np.diff(sm.astype(np.float64)), np.diff(seg.astype(np.int64))
            assert np.all((ds < 0) | ((ds == 0) & (di > 0)))
            if psz[p]:
                assert np.array_equal(pmax[p], x[:, seg].max(axis=1))
                assert np.array_equal(pmin[p], x[:, seg].min(axis=1))
        assert 0.5 * 4096 < psz.mean() < 1.5 * 4096


@pytest.mark.parametrize("distr", [0, 1, 2, 3])
def test_pta_parity(gpu, distr):
    n, d = 1_000_003, 8
    x = topk_amd.generate(n, d, distr, 40 + distr)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        lay = _layout(eng, x)
        for k in (1, 16, 128, 1024):
            _check_pta(eng, x, n, k, np.ones(d, np.float32), list(range(8)), lay)
        for m in (2, 5):
            _check_pta(eng, x, n, 64, RNG.uniform(0.1, 1, d).astype(np.float32), list(range(d - m, d)), lay)
        st = _check_pta(eng, x, n, 32, RNG.uniform(0.1, 1, d).astype(np.float32), [6, 1, 3], lay)
        assert st["partitions_total"] == lay[3].size


def test_pta_prunes_partitions(gpu):
    n, d = 2_000_000, 8
    x = topk_amd.generate(n, d, topk_amd.MIXED, 3)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        st = _check_pta(eng, x, n, 64, RNG.uniform(0.1, 1, d).astype(np.float32), list(range(8)))
        assert st["partitions_scored"] < 0.2 * st["partitions_total"]


def test_pta_negative_weights_duplicates_ties(gpu):
    n, d = 400_000, 6
    x = (RNG.integers(0, 3, size=(d, n)) / 2.0).astype(np.float32)  # heavy ties
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        w = np.array([0.5, -1.25, 3.0, -0.1, 0.0, 2.0], np.float32)
        _check_pta(eng, x, n, 100, w, [1, 3, 0])
        _check_pta(eng, x, n, 100, w, [2, 2, 5, 4])
        _check_pta(eng, x, n, 500, np.ones(d, np.float32), list(range(d)))


@pytest.mark.parametrize("n", [1, 5, 1000, 4096, 5000, 70_000])
def test_pta_small_sizes(gpu, n):
    d = 3
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, n)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        for k in (1, 16, 64):
            _check_pta(eng, x, n, k, np.ones(d, np.float32), [0, 2])


def test_pta_partition_size_and_state(gpu):
    n, d = 200_000, 4
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 2)
    with topk_amd.Engine(x) as eng:
        with pytest.raises(topk_amd.TopkError) as e:
            eng.pta(8, np.ones(d, np.float32), [0])
        assert e.value.status == 6
        with pytest.raises(topk_amd.TopkError):
            eng.set_partition_size(1000)
        eng.set_partition_size(2048)
        eng.init(topk_amd.TOPK_INIT_PTA)
        _, _, psz, _, _ = eng.pta_layout()
        assert 0.5 * 2048 < psz.mean() < 1.5 * 2048
        _check_pta(eng, x, n, 50, RNG.uniform(0.1, 1, d).astype(np.float32), [3, 0, 1])


@pytest.mark.parametrize("n", [100_000, 100_003])
def test_normalize_matches_oracle(gpu, n):
    d = 5
    x = (topk_amd.generate(n, d, topk_amd.INDEPENDENT, 5) * 7.0 - 2.0).astype(np.float32)
    x[3] = -np.abs(x[3]) - 1.0  # an all-negative column keeps max 0 (common/common.h:36)
    x[0, n - 1], x[1, n - 1] = 50.0, -50.0  # column extremes in the last object (past the last whole quad)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_NORMALIZE | topk_amd.TOPK_INIT_PTA)
        y = eng.read_columns()
        ref = oracle.normalize(x, n)
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
        ids, sc, _ = eng.fte(20, np.ones(d, np.float32), [0, 1, 2, 3, 4])
        rid, rsc, _ = oracle.fte_topk(ref, n, 20, np.ones(d, np.float32), [0, 1, 2, 3, 4])
        assert np.array_equal(ids, rid) and np.array_equal(sc, rsc)
        _check_pta(eng, ref, n, 20, np.ones(d, np.float32), [0, 1, 2, 3, 4])


@pytest.mark.parametrize("case", ["constant", "range_overflow"])
def test_normalize_failure_leaves_store_untouched(gpu, case):
    n, d = 1000, 3
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 9)
    if case == "constant":
        x[1] = 0.75  # max - min == 0: every value would be NaN
    else:
        x[2, 5], x[2, 7] = 3.0e38, -3.0e38  # max - min overflows to infinity
    with topk_amd.Engine(x) as eng:
        with pytest.raises(topk_amd.TopkError) as e:
            eng.init(topk_amd.TOPK_INIT_NORMALIZE)
        assert e.value.status == 5
        # nothing was written: the store and later queries are the originals
        assert np.array_equal(eng.read_columns().view(np.uint32), x.view(np.uint32))
        w = np.ones(d, np.float32)
        ids, sc, _ = eng.fte(10, w, [0, 1])
        rid, rsc, _ = oracle.fte_topk(x, n, 10, w, [0, 1])
        assert np.array_equal(ids, rid) and np.array_equal(sc, rsc)


def test_normalize_zero_sign_follows_first_zero(gpu):
    # The column minimum is zero with both signs present: the reference's
    # sequential std::min keeps the first zero it meets (common/common.h:29-60),
    # which decides the sign of the normalised zeros.
    n = 70_001
    x = topk_amd.generate(n, 2, topk_amd.INDEPENDENT, 12) + np.float32(0.5)
    x[0, 40_000], x[0, 50_000] = 0.0, -0.0   # +0 first
    x[1, 123], x[1, 60_000] = -0.0, 0.0      # -0 first
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_NORMALIZE)
        y = eng.read_columns()
    ref = oracle.normalize(x, n)
    assert np.signbit(ref[0, 50_000]) and not np.signbit(ref[1, 123])
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_reorder_attr_matches_oracle(gpu):
    n, d = 123_457, 8
    x = topk_amd.generate(n, d, topk_amd.CORRELATED, 6)
    with topk_amd.Engine(x) as eng:
        eng.reorder_attr()
        y = eng.read_columns()
        ref = oracle.reorder_attr(x, n)
        assert np.array_equal(y, ref)
        assert oracle.check_order(y, n) == n
        ids, sc, _ = eng.fte(16, np.ones(d, np.float32), list(range(8)))
        rid, rsc, _ = oracle.fte_topk(ref, n, 16, np.ones(d, np.float32), list(range(8)))
        assert np.array_equal(ids, rid) and np.array_equal(sc, rsc)


def test_small_layout_seeds_from_several_partitions(gpu):
    """N = 2^20 (256 partitions) with a 4-of-8 attribute query: the seed
    ranks the partitions' own bounds, so tau leaves a small candidate set
    (seeding from the one partition the block maxima rank first made this
    query's final select take 16 ms)."""
    n, d = 1 << 20, 8
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 0)
    w = np.ones(d, np.float32)
    with topk_amd.Engine(x) as eng:
        eng.init(topk_amd.TOPK_INIT_PTA)
        got = eng.pta(16, w, [4, 5, 6, 7])
        st = eng.stats()
        ref = oracle.fte_topk(x, n, 16, w, [4, 5, 6, 7])
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert st["candidates"] < 2000, st
