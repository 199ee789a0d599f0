"""GPU parity of the FTE path (topk_fte / topk_fte_async) against the CPU
oracle: ids and scores must match bit for bit."""
import numpy as np
import pytest

import oracle
import topk_amd

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(99)


def _check(eng, x, n, k, w, attr):
    ids, sc, thr = eng.fte(k, w, attr)
    rid, rsc, rthr = oracle.fte_topk(x, n, k, w, attr)
    assert np.array_equal(ids, rid), (n, k, list(attr))
    assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32)), (n, k, list(attr))
    assert thr == rthr
    return ids, sc


@pytest.mark.parametrize("distr", [0, 1, 2, 3])
def test_fte_parity_distributions(gpu, distr):
    n, d = 300_007, 8
    x = topk_amd.generate(n, d, distr, 11 + distr)
    with topk_amd.Engine(x) as eng:
        for m in range(1, d + 1):
            attr = list(range(d - m, d))  # bench_cpu.h QM=0 rows: last m attributes
            for k in (1, 16, 64, 128):
                _check(eng, x, n, k, np.ones(d, np.float32), attr)
        w = RNG.uniform(0.1, 1.0, d).astype(np.float32)
        _check(eng, x, n, 64, w, [5, 1, 7, 0, 3])


@pytest.mark.parametrize("n", [1, 3, 15, 16, 17, 63, 64, 65, 1000, 4095, 4096, 4097, 8193, 1_048_577])
def test_fte_parity_ragged_sizes(gpu, n):
    d = 4
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, n)
    with topk_amd.Engine(x) as eng:
        for k in (1, 16, 64):
            _check(eng, x, n, k, np.ones(d, np.float32), [1, 2, 3])


def test_fte_k_larger_than_n_sentinels(gpu):
    n, d = 10, 2
    x = topk_amd.generate(n, d, 1, 5)
    with topk_amd.Engine(x) as eng:
        ids = np.empty(16, np.uint64)
        sc = np.empty(16, np.float32)
        import ctypes
        thr = np.zeros(1, np.float32)
        w = np.ones(32, np.float32)
        at = np.array([0, 1], np.uint8)
        rc = topk_amd.lib().topk_fte(eng.handle, 16, 2, w.ctypes.data, at.ctypes.data, ids.ctypes.data,
                                     sc.ctypes.data, thr.ctypes.data)
        assert rc == 0
        rid, rsc, _ = oracle.fte_topk(x, n, 16, w, [0, 1])
        assert np.array_equal(ids[:10], rid) and np.all(ids[10:] == np.uint64(2 ** 64 - 1))
        assert np.all(np.isneginf(sc[10:]))
        assert thr[0] == rsc[-1]


def test_fte_large_k(gpu):
    n, d = 200_000, 8
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 3)
    with topk_amd.Engine(x) as eng:
        for k in (256, 512, 1000, 1024):
            _check(eng, x, n, k, RNG.uniform(0.1, 1, d).astype(np.float32), list(range(d)))


def test_fte_ties_all_equal(gpu):
    # every score equal: the top-k are the k lowest ids
    n, d = 100_000, 3
    x = np.full((d, n), 0.5, np.float32)
    with topk_amd.Engine(x) as eng:
        ids, sc, thr = eng.fte(64, np.ones(d, np.float32), [0, 1, 2])
        assert ids.tolist() == list(range(64)) and np.all(sc == 1.5)


def test_fte_ties_quantised(gpu):
    n, d = 500_000, 4
    x = (RNG.integers(0, 3, size=(d, n)) / 2.0).astype(np.float32)
    with topk_amd.Engine(x) as eng:
        for k in (16, 128, 1024):
            _check(eng, x, n, k, np.ones(d, np.float32), [0, 1, 2, 3])


def test_fte_correlated_ties_at_max(gpu):
    # correlated data clips to 1.0: many objects score exactly m
    n, d = 2_000_000, 8
    x = topk_amd.generate(n, d, topk_amd.CORRELATED, 0)
    with topk_amd.Engine(x) as eng:
        for m in (2, 4, 8):
            ids, sc = _check(eng, x, n, 128, np.ones(d, np.float32), list(range(d - m, d)))
            if m == 2:
                assert np.all(sc == 2.0)


def test_fte_negative_weights_and_repeats(gpu):
    n, d = 123_457, 6
    x = topk_amd.generate(n, d, topk_amd.MIXED, 4)
    with topk_amd.Engine(x) as eng:
        w = np.array([0.5, -1.25, 3.0, -0.1, 0.0, 2.0], np.float32)
        _check(eng, x, n, 32, w, [1, 3, 0])
        _check(eng, x, n, 32, w, [2, 2, 5, 4])
        _check(eng, x, n, 32, -np.ones(d, np.float32), [0, 1, 2, 3, 4, 5])


def test_fte_wide_queries(gpu):
    n, d = 70_001, 32
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 8)
    with topk_amd.Engine(x) as eng:
        w = RNG.uniform(0.1, 1, d).astype(np.float32)
        for m in (9, 16, 17, 24, 32):
            attr = RNG.permutation(d)[:m]
            _check(eng, x, n, 64, w, attr)


def test_fte_rowmajor_ingest(gpu):
    n, d = 33_333, 5
    x = topk_amd.generate(n, d, topk_amd.CORRELATED, 9)
    with topk_amd.Engine(np.ascontiguousarray(x.T), rowmajor=True) as eng:
        assert np.array_equal(eng.read_columns(), x)
        _check(eng, x, n, 16, np.ones(d, np.float32), [0, 4])


def test_generated_store_matches_host_generator(gpu):
    for distr in range(4):
        n, d = 77_777, 7
        with topk_amd.Engine.generated(n, d, distr, 42) as eng:
            assert np.array_equal(eng.read_columns(), topk_amd.generate(n, d, distr, 42))


def test_nonfinite_rejected(gpu):
    x = np.ones((2, 100), np.float32)
    x[1, 50] = np.nan
    with pytest.raises(topk_amd.TopkError) as e:
        topk_amd.Engine(x)
    assert e.value.status == 5


def test_fte_overflow_falls_back_exactly(gpu):
    # scores rise with the id, so the exact-prefix threshold is weak and more
    # objects pass the stream filter than the candidate slots (2^21) and the
    # spill region (2^21) hold together
    n = 8_000_000
    x = (np.arange(n, dtype=np.float64) / n).astype(np.float32)[None, :]
    with topk_amd.Engine(x) as eng:
        _check(eng, x, n, 64, np.ones(1, np.float32), [0])
        assert eng.stats()["fallback"] == 0  # the spread sample copes
        eng.set_tuning("fte_force_exact", 1)
        _check(eng, x, n, 64, np.ones(1, np.float32), [0])
        assert eng.stats()["fallback"] == 1
        _check(eng, x, n, 64, -np.ones(1, np.float32), [0])
        assert eng.stats()["fallback"] == 0


def test_fte_async_matches_sync(gpu):
    import torch
    n, d = 1_000_000, 8
    with topk_amd.Engine.generated(n, d, topk_amd.INDEPENDENT, 1) as eng:
        k = 64
        ids_d = torch.empty(4 * k, dtype=torch.int64, device="cuda")
        sc_d = torch.empty(4 * k, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        qs = [(np.ones(d, np.float32), list(range(8 - m, 8))) for m in (2, 4, 6, 8)]
        for i, (w, at) in enumerate(qs):
            eng.fte_async(k, w, at, ids_d.data_ptr() + 8 * k * i, sc_d.data_ptr() + 4 * k * i)
        eng.sync()
        ids_h = ids_d.cpu().numpy().view(np.uint64)
        sc_h = sc_d.cpu().numpy()
        for i, (w, at) in enumerate(qs):
            ids, sc, _ = eng.fte(k, w, at)
            assert np.array_equal(ids_h[i * k:(i + 1) * k], ids)
            assert np.array_equal(sc_h[i * k:(i + 1) * k], sc)


@pytest.mark.parametrize("distr,seed", [(topk_amd.INDEPENDENT, 0), (topk_amd.CORRELATED, 1),
                                        (topk_amd.ANTICORRELATED, 2)])
def test_fte_full_size_sweep(gpu, distr, seed):
    """The plain-FTE benchmark sweep at its full size: N = 2^26, d = 8, the
    last m = 2..8 attributes (validation/bench_cpu.h:41-44, QM=0 rows),
    weights 1.0, k in {16, 32, 64, 128}, every point against the oracle's
    SIMD+threads TPAc restatement (cpu/TPAc.h:256-358) on the same bits."""
    n, d = 1 << 26, 8
    x = topk_amd.generate(n, d, distr, seed, threads=16)
    w = np.ones(d, np.float32)
    with topk_amd.Engine.generated(n, d, distr, seed) as eng:
        for m in range(2, d + 1):
            attr = list(range(d - m, d))
            for k in (16, 32, 64, 128):
                ids, sc, thr = eng.fte(k, w, attr)
                rid, rsc, rthr = oracle.tpac_threads(x, n, k, w, attr, 16)
                assert np.array_equal(ids, rid), (distr, m, k)
                assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32)), (distr, m, k)
                assert np.float32(thr).tobytes() == np.float32(rthr).tobytes(), (distr, m, k)
    del x


def test_fte_exact_seed_mode_parity(gpu):
    n, d = 2_500_000, 8
    x = topk_amd.generate(n, d, topk_amd.MIXED, 21)
    with topk_amd.Engine(x) as eng:
        eng.set_tuning("fte_force_exact", 1)
        for k in (1, 64, 256):
            _check(eng, x, n, k, RNG.uniform(0.1, 1, d).astype(np.float32), [0, 2, 4, 6, 7])
        eng.set_tuning("fte_force_exact", 0)
        _check(eng, x, n, 64, np.ones(d, np.float32), [0, 2, 4, 6, 7])


def test_fte_sample_guess_too_high_retries(gpu):
    # the best objects sit inside the first sample chunk, so the sample's
    # 4th-best key admits fewer than k candidates and the exact path reruns
    n, d = 4_000_000, 2
    x = (RNG.random((d, n), dtype=np.float32) * 0.5).astype(np.float32)
    x[:, :200] = 0.9
    x[0, :200] += np.arange(200, dtype=np.float32) * 1e-4
    with topk_amd.Engine(x) as eng:
        _check(eng, x, n, 64, np.ones(d, np.float32), [0, 1])
        assert eng.stats()["fallback"] == 1


@pytest.mark.parametrize("vpt", [1, 2, 4, 8])
def test_fte_stream_shapes(gpu, vpt):
    n, d = 3_000_001, 8
    x = topk_amd.generate(n, d, topk_amd.ANTICORRELATED, 5)
    with topk_amd.Engine(x) as eng:
        eng.set_tuning("stream_vpt", vpt)
        for grid in (1, 7, 300, 0):
            eng.set_tuning("stream_grid", grid)
            _check(eng, x, n, 32, np.ones(d, np.float32), list(range(8)))
            eng.set_tuning("fte_force_exact", 1)
            _check(eng, x, n, 32, np.ones(d, np.float32), [7, 3])
            eng.set_tuning("fte_force_exact", 0)


def test_golden_fixtures_all_paths(gpu):
    """FTE, PTA and batch against the pure-numpy fixtures of
    tests/golden/make_golden.py (data generated on the device)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
    n, d = 1 << 16, 8
    cases = ((1, 0, 16, np.ones(d, np.float32), np.array([4, 5, 6, 7], np.uint8), "c1"),
             (0, 2, 64, g["cw_weights"], g["cw_attr"], "cw"))
    for distr, seed, k, w, attr, tag in cases:
        with topk_amd.Engine.generated(n, d, distr, seed) as eng:
            eng.init(topk_amd.TOPK_INIT_PTA)
            outs = [eng.fte(k, w, attr), eng.pta(k, w, attr), eng.batch([(k, w, attr)])[0]]
            for ids, sc, thr in outs:
                assert np.array_equal(ids, g[f"{tag}_ids"]), tag
                assert np.array_equal(sc.view(np.uint32), g[f"{tag}_scores"]), tag


@pytest.mark.parametrize("n", [100_003, 3_000_017])
def test_fte_real_data_scores_match_oracle(gpu, n):
    """topk_set_tuning "score_divide": the running score is divided by qq
    after every term (the reference's LD == 2 build, cpu/TPAc.h:143-146,
    237-240), bit-exact against the oracle's restatement; batches in that
    mode answer every query through the FTE path; off again, plain sums."""
    d = 8
    x = topk_amd.generate(n, d, topk_amd.INDEPENDENT, 9)
    rng = np.random.default_rng(5)
    cases = [(16, 2), (64, 3), (100, 5), (128, 8), (7, 1)]
    qs = []
    with topk_amd.Engine(x) as eng:
        eng.set_tuning("score_divide", 1)
        for k, qq in cases:
            attr = [int(a) for a in rng.permutation(d)[:qq]]
            w = rng.uniform(0.1, 1.0, d).astype(np.float32)
            qs.append((k, w, attr))
            ids, sc, thr = eng.fte(k, w, attr)
            rid, rsc, rthr = oracle.fte_topk(x, n, k, w, attr, divide=True)
            assert np.array_equal(ids, rid), (n, k, qq)
            assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32)), (n, k, qq)
            assert np.float32(thr).tobytes() == np.float32(rthr).tobytes()
        for (k, w, attr), (gi, gs, gt) in zip(qs, eng.batch(qs)):
            rid, rsc, rthr = oracle.fte_topk(x, n, k, w, attr, divide=True)
            assert np.array_equal(gi, rid) and np.array_equal(gs.view(np.uint32), rsc.view(np.uint32))
            assert np.float32(gt).tobytes() == np.float32(rthr).tobytes()
        assert eng.stats()["fte_fallback"] == len(qs)
        with pytest.raises(topk_amd.TopkError):
            eng.batch_autotune(qs)
        eng.set_tuning("score_divide", 0)
        k, w, attr = qs[3]
        ids, sc, _ = eng.fte(k, w, attr)
        rid, rsc, _ = oracle.fte_topk(x, n, k, w, attr)
        assert np.array_equal(ids, rid) and np.array_equal(sc, rsc)
