"""Pins the CPU oracle (oracle/topk_oracle.c) without trusting it.

Every check is a closed form, a hand-derived known answer, a brute-force
computation by a different method, or an invariant -- chosen so that a
plausible mistake (summation order, an FMA, a dropped term, a wrong
tie-break, an off-by-one in a radix digit) fails one of them.
"""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1234)


def _dyadic_case(n, d, qq, rng):
    """x = a/256, w = b/16 with small integers: every product and partial sum
    is exact in binary32, so score == sum(a*b)/4096 exactly."""
    a = rng.integers(0, 256, size=(d, n))
    b = rng.integers(1, 17, size=d)
    attr = rng.permutation(d)[:qq].astype(np.uint8)
    x = (a / 256.0).astype(np.float32)
    w = (b / 16.0).astype(np.float32)
    exact = (a[attr, :] * b[attr, None]).sum(axis=0)  # integer numerator over 4096
    return x, w, attr, exact


def test_keys_order():
    vals = np.array([-np.inf, -3e38, -1.5, -1e-30, 0.0, 1e-38, 1e-30, 0.5, 1.0, 7.0, 3e38, np.inf],
                    dtype=np.float32)
    ks = [oracle.okey(v) for v in vals]
    assert ks == sorted(ks) and len(set(ks)) == len(ks)
    # (score desc, id asc): higher score wins; on equal score the lower id wins
    assert oracle.key(2.0, 10) > oracle.key(1.0, 0)
    assert oracle.key(1.0, 3) > oracle.key(1.0, 4)
    assert oracle.key(-np.inf, 0xFFFFFFFE) > 0


def test_score_closed_form_dyadic():
    for qq in (1, 2, 5, 8):
        x, w, attr, exact = _dyadic_case(3000, 8, qq, RNG)
        s = oracle.score(x, 3000, w, attr)
        assert np.array_equal(s.astype(np.float64) * 4096.0, exact.astype(np.float64))


def test_score_summation_order_known_answer():
    # products 1, 2^-24, 2^-24: ((1 + 2^-24) + 2^-24) = 1 (ties to even)
    # but ((2^-24 + 2^-24) + 1) = 1 + 2^-23.
    x = np.array([[1.0], [2.0 ** -24], [2.0 ** -24]], dtype=np.float32)
    w = np.ones(3, dtype=np.float32)
    assert oracle.score(x, 1, w, [0, 1, 2])[0] == np.float32(1.0)
    assert oracle.score(x, 1, w, [2, 1, 0])[0] == np.float32(1.0 + 2.0 ** -23)


def test_score_no_fma_known_answer():
    # attr order [1, 0]: acc = -1, then acc + fl((1+2^-12)^2).  Separately
    # rounded: fl((1+2^-12)^2) = 1 + 2^-11 (the 2^-24 term is a tie to even),
    # so the score is 2^-11; a fused multiply-add would give 2^-11 + 2^-24.
    x = np.array([[1.0 + 2.0 ** -12], [1.0]], dtype=np.float32)
    w = np.array([1.0 + 2.0 ** -12, -1.0], dtype=np.float32)
    assert oracle.score(x, 1, w, [1, 0])[0] == np.float32(2.0 ** -11)


def test_weights_indexed_by_attribute_id():
    # cpu/TPAc.h:135: weight = weights[attr[m]], not weights[m]
    x = np.array([[1.0], [10.0], [100.0]], dtype=np.float32)
    w = np.array([1.0, 2.0, 3.0], dtype=np.float32)
    assert oracle.score(x, 1, w, [2, 0])[0] == np.float32(301.0)


@pytest.mark.parametrize("k", [1, 7, 16, 64])
def test_fte_topk_bruteforce_with_ties(k):
    # quantised dyadic data -> many exact score ties; the brute-force answer
    # is a full lexicographic sort by (-score, id) of exact integer scores
    n = 2000
    a = RNG.integers(0, 4, size=(4, n))
    x = (a / 4.0).astype(np.float32)
    w = np.ones(4, dtype=np.float32)
    exact = a[[1, 3], :].sum(axis=0)
    order = np.lexsort((np.arange(n), -exact))[:k]
    ids, sc, thr = oracle.fte_topk(x, n, k, w, [1, 3])
    assert np.array_equal(ids, order.astype(np.uint64))
    assert np.array_equal(sc * 4.0, exact[order].astype(np.float32))
    assert thr == sc[-1]


def test_score_divide_powers_of_two_closed_form():
    """Real-data scores (LD == 2): with qq = 2 or 4, w = 1 and x = a/256,
    every division is by a power of two and every step is exact, so term j
    (0-based) of qq ends up divided qq - j times:
    score = sum_j a_j / 256 / qq^(qq - j)."""
    rng = np.random.default_rng(7)
    n, d = 4000, 8
    a = rng.integers(0, 256, size=(d, n))
    x = (a / 256.0).astype(np.float32)
    w = np.ones(d, np.float32)
    for qq in (2, 4):
        attr = rng.permutation(d)[:qq].astype(np.uint8)
        got = oracle.score(x, n, w, attr, divide=True)
        want = sum(a[attr[j]] / 256.0 / float(qq) ** (qq - j) for j in range(qq))
        assert np.array_equal(got.astype(np.float64), want)
    # qq = 1 divides by one: the plain score
    assert np.array_equal(oracle.score(x, n, w, [3], divide=True), oracle.score(x, n, w, [3]))


def test_score_divide_matches_float32_steps():
    """Other qq: numpy's binary32 mul, add and divide (each correctly
    rounded) applied term by term in attribute order."""
    rng = np.random.default_rng(8)
    n, d = 3000, 8
    x = rng.random((d, n), dtype=np.float32)
    w = rng.uniform(0.1, 1.0, d).astype(np.float32)
    for qq in (3, 5, 7, 8):
        attr = rng.permutation(d)[:qq].astype(np.uint8)
        acc = np.zeros(n, np.float32)
        q = np.float32(qq)
        for j in range(qq):
            acc = (acc + (x[attr[j]] * w[attr[j]]).astype(np.float32)).astype(np.float32)
            acc = (acc / q).astype(np.float32)
        got = oracle.score(x, n, w, attr, divide=True)
        assert np.array_equal(got.view(np.uint32), acc.view(np.uint32)), qq
        ids, sc, thr = oracle.fte_topk(x, n, 25, w, attr, divide=True)
        order = np.lexsort((np.arange(n), -acc.astype(np.float64)))[:25]
        assert np.array_equal(ids, order.astype(np.uint64)) and np.array_equal(sc, acc[order])
        assert thr == acc[order[-1]]


def test_fte_topk_n_smaller_than_k():
    x = (RNG.integers(0, 100, size=(2, 5)) / 128.0).astype(np.float32)
    ids, sc, thr = oracle.fte_topk(x, 5, 16, np.ones(2, np.float32), [0, 1])
    assert len(ids) == 5 and sorted(ids.tolist()) == list(range(5))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_tpac_threads_matches_scalar(threads):
    n = 50_001
    x = RNG.random((8, n), dtype=np.float32)
    w = RNG.uniform(0.1, 1.0, 8).astype(np.float32)
    for attr, k in (([4, 5, 6, 7], 16), ([7, 0, 3], 100), (list(range(8)), 1)):
        a = oracle.fte_topk(x, n, k, w, attr)
        b = oracle.tpac_threads(x, n, k, w, attr, threads)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("threads", [1, 5])
def test_fte_topk_wide_matches_32bit_below_2e32(threads):
    """Below 2^32 the 128-bit-key restatement ranks as the 64-bit-key one."""
    n = 30_011
    x = (RNG.integers(0, 64, size=(4, n)) / 64.0).astype(np.float32)  # many ties
    w = RNG.uniform(0.1, 1.0, 4).astype(np.float32)
    for attr, k in (([0, 1, 2, 3], 50), ([2], 300), ([3, 1], 1)):
        a = oracle.fte_topk(x, n, k, w, attr)
        b = oracle.fte_topk_wide(x, n, k, w, attr, threads)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


@pytest.mark.parametrize("threads", [1, 4])
def test_fte_topk_wide_ids_across_2e32_bruteforce(threads):
    """Ids id_base + i crossing 2^32 (and 2^33): ranked by a lexicographic
    sort on (-score, id) of numpy-computed scores; quantised values make
    ties straddle the boundary, where the lower 64-bit id must win."""
    n, d = 4_000, 2
    for base in ((1 << 32) - 1_500, (1 << 33) - 7):
        x = (RNG.integers(0, 8, size=(d, n)) / 8.0).astype(np.float32)
        w = np.array([1.0, 0.5], np.float32)
        sc = (np.float32(0.0) + x[0] * w[0]).astype(np.float32) + (x[1] * w[1]).astype(np.float32)
        ids = base + np.arange(n, dtype=np.uint64)
        order = np.lexsort((ids, -sc.astype(np.float64)))
        for k in (1, 37, 600):
            gi, gs, gt = oracle.fte_topk_wide(x, n, k, w, [0, 1], threads, id_base=base)
            assert np.array_equal(gi, ids[order[:k]])
            assert np.array_equal(gs, sc[order[:k]]) and gt == sc[order[k - 1]]
        # a tie across 2^32: the object below the boundary ranks first
        if base < (1 << 32):
            lo, hi = (1 << 32) - 1 - base, (1 << 32) + 5 - base
            x2 = np.zeros((d, n), np.float32)
            x2[0, hi] = 1.0
            x2[0, lo] = 1.0
            gi, gs, _ = oracle.fte_topk_wide(x2, n, 2, w, [0, 1], threads, id_base=base)
            assert gi.tolist() == [(1 << 32) - 1, (1 << 32) + 5] and gs.tolist() == [1.0, 1.0]


def test_mq_threads_matches_single():
    n = 20_000
    x = RNG.random((6, n), dtype=np.float32)
    ks = np.array([5, 16, 33], dtype=np.uint32)
    qqs = np.array([2, 6, 3], dtype=np.uint32)
    attrs = np.array([[4, 5, 0, 0, 0, 0], [0, 1, 2, 3, 4, 5], [5, 1, 3, 0, 0, 0]], dtype=np.uint8)
    ws = RNG.uniform(0.1, 1, (3, 6)).astype(np.float32)
    res = oracle.mq_threads(x, n, ks, qqs, ws, attrs, 3)
    for q in range(3):
        ids, sc, _ = oracle.fte_topk(x, n, int(ks[q]), ws[q], attrs[q, :qqs[q]])
        assert np.array_equal(res[q][0], ids) and np.array_equal(res[q][1], sc)


def test_radix_select_matches_sort():
    for n in (1, 2, 17, 5000):
        keys = RNG.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[0]  # duplicates
        s = np.sort(keys)
        for r in {0, n // 2, n - 1}:
            assert oracle.radix_select_u32(keys, r) == s[r]
        k64 = RNG.integers(0, 2 ** 63, size=n, dtype=np.uint64) * 2 + 1
        s64 = np.sort(k64)
        for r in {0, n // 3, n - 1}:
            assert oracle.radix_select_u64(k64, r) == s64[r]


def test_radix_select_float_threshold():
    # gpu/GPA.h:183-187 known-answer method: select(n-k) on float bit
    # patterns of non-negative scores == k-th largest by sort
    f = RNG.random(10_000, dtype=np.float32)
    bits = f.view(np.uint32)
    k = 64
    got = np.uint32(oracle.radix_select_u32(bits, f.size - k)).view(np.float32)
    assert got == np.sort(f)[-k]


def test_normalize_closed_form():
    x = np.array([[2.0, 4.0, 6.0], [-4.0, -2.0, -3.0]], dtype=np.float32)
    y = oracle.normalize(x, 3)
    assert np.array_equal(y[0], np.array([0.0, 0.5, 1.0], np.float32))
    # max starts at 0 (common/common.h:36): an all-negative column keeps max 0
    assert np.array_equal(y[1], np.array([0.0, 0.5, 0.25], np.float32))


def test_normalize_zero_sign_known_answer():
    # std::min(mmin, v) keeps mmin when v == mmin, so the first zero met fixes
    # the minimum's sign: (-0 - +0) / 1 = -0, while (-0 - -0) = +0.
    x = np.array([[0.0, -0.0, 1.0], [-0.0, 0.0, 2.0]], dtype=np.float32)
    y = oracle.normalize(x, 3)
    assert np.signbit(y[0, 1]) and not np.signbit(y[0, 0])
    assert not np.signbit(y[1, 0]) and not np.signbit(y[1, 1])
    assert np.array_equal(y[:, 2], np.ones(2, np.float32))


def test_reorder_attr_invariants():
    n, d = 3000, 8
    x = RNG.random((d, n), dtype=np.float32)
    x[:, :10] = 0.5  # ties
    y = oracle.reorder_attr(x, n)
    assert oracle.check_order(y, n) == n
    assert oracle.check_order(x, n) < n
    assert np.array_equal(np.sort(x, axis=0), np.sort(y, axis=0))
    assert np.array_equal(y[0], x.max(axis=0))
    assert np.array_equal(y[-1], x.min(axis=0))


def test_bta_blocks_match_sort():
    n, blk, k = 10_000, 4096, 32
    x = RNG.random((3, n), dtype=np.float32)
    w = np.array([0.3, 1.0, 0.7], dtype=np.float32)
    out = oracle.bta_blocks(x, n, w, [2, 0, 1], blk, k)
    s = oracle.score(x, n, w, [2, 0, 1])
    for b in range(out.shape[0]):
        lo, hi = b * blk, min(n, (b + 1) * blk)
        order = np.lexsort((np.arange(lo, hi), -s[lo:hi]))[:k] + lo
        keys = [oracle.key(s[i], i) for i in order]
        assert out[b].tolist() == keys


def test_gpa_step_known_answer():
    ids = np.array([0, 1, 2, 3], dtype=np.uint32)
    score = np.array([1.0, 0.5, 0.25, 0.0], dtype=np.float32)
    colC = np.array([0.5, 0.25, 0.5, 0.125], dtype=np.float32)
    colN = np.array([1.0, 2.0, 3.0, 4.0], dtype=np.float32)
    # upper bounds with suffix_len 2: 2.0, 1.0, 1.25, 0.25; keep >= 1.0
    i2, s2 = oracle.gpa_step(ids, score, colC, colN, 1.0, 2.0)
    assert i2.tolist() == [0, 1, 2] and s2.tolist() == [2.0, 2.5, 3.25]


def _random_layout(x, n, psize, rng, cluster):
    """A partition layout like the engine's: partitions of psize objects
    (last one short), each padded to a multiple of 256 positions."""
    d = x.shape[0]
    if cluster:
        perm = np.lexsort(tuple((x[m] * 4).astype(int) for m in range(d)))
    else:
        perm = rng.permutation(n)
    npart = (n + psize - 1) // psize
    psz = np.array([min(psize, n - p * psize) for p in range(npart)], np.uint32)
    pad = ((psz + 255) // 256) * 256
    poff = np.concatenate([[0], np.cumsum(pad)]).astype(np.uint32)
    npos = int(poff[-1])
    pdata = np.zeros((d, npos), np.float32)
    pid = np.full(npos, 0xFFFFFFFF, np.uint32)
    pmax = np.zeros((npart, d), np.float32)
    pmin = np.zeros((npart, d), np.float32)
    for p in range(npart):
        ids = perm[p * psize:p * psize + psz[p]]
        pdata[:, poff[p]:poff[p] + psz[p]] = x[:, ids]
        pid[poff[p]:poff[p] + psz[p]] = ids
        pmax[p] = x[:, ids].max(axis=1)
        pmin[p] = x[:, ids].min(axis=1)
    return pdata, pid, poff, psz, pmax, pmin


@pytest.mark.parametrize("cluster", [False, True])
def test_pta_equals_fte(cluster):
    n, d, psize = 40_000, 4, 1000
    x = RNG.random((d, n), dtype=np.float32)
    pdata, pid, poff, psz, pmax, pmin = _random_layout(x, n, psize, RNG, cluster)
    for w, attr, k in ((np.ones(d, np.float32), [0, 1, 2, 3], 16),
                       (RNG.uniform(0.1, 1, d).astype(np.float32), [3, 1], 50),
                       (np.array([1.0, -0.5, 0.25, 2.0], np.float32), [1, 0, 3], 8)):
        ref = oracle.fte_topk(x, n, k, w, attr)
        got = oracle.pta_topk(pdata, pid, poff, psz, pmax, pmin, k, w, attr)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        got_t = oracle.pta_topk(pdata, pid, poff, psz, pmax, pmin, k, w, attr, threads=4)
        assert np.array_equal(got_t[0], ref[0]) and np.array_equal(got_t[1], ref[1])
        if cluster:
            assert got[3] < pmax.shape[0]  # some partitions were skipped


def test_pta_bound_is_upper_bound():
    n, d, psize = 8192, 3, 1024
    x = RNG.random((d, n), dtype=np.float32)
    pdata, pid, poff, psz, pmax, pmin = _random_layout(x, n, psize, RNG, True)
    w = np.array([0.7, -1.3, 0.2], np.float32)
    attr = [2, 1, 0]
    s = oracle.score(pdata, pdata.shape[1], w, attr)
    for p in range(pmax.shape[0]):
        b = oracle.pta_bound(pmax, pmin, d, p, w, attr)
        assert s[poff[p]:poff[p] + psz[p]].max() <= b


def test_mlp_gradient_finite_differences():
    nparam = oracle.mlp_nparams()
    assert nparam == 96 * 64 + 64 + 64 * 64 + 64 + 64 + 1
    p = RNG.normal(0, 0.2, nparam)
    x = RNG.normal(0, 1, (5, 96))
    y = RNG.normal(0, 1, 5)
    loss, g = oracle.mlp_backward(p, x, y)
    pred = oracle.mlp_forward(p, x)
    assert abs(loss - np.mean((pred - y) ** 2)) < 1e-12
    idx = RNG.choice(nparam, 60, replace=False).tolist() + [nparam - 1, 0, 96 * 64]
    for i in idx:
        e = 1e-6
        pp = p.copy(); pp[i] += e
        pm = p.copy(); pm[i] -= e
        fd = (np.mean((oracle.mlp_forward(pp, x) - y) ** 2) - np.mean((oracle.mlp_forward(pm, x) - y) ** 2)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_forward_closed_form():
    p = np.zeros(oracle.mlp_nparams())
    p[-1] = 3.25  # b3 only -> constant output
    assert np.all(oracle.mlp_forward(p, RNG.normal(0, 1, (4, 96))) == 3.25)
    # one path: W1[0][5]=2, W2[0][0]=3, W3[0]=0.5 -> out = 0.5*relu(3*relu(2*x5))
    p = np.zeros(oracle.mlp_nparams())
    p[5] = 2.0
    p[96 * 64 + 64] = 3.0
    p[96 * 64 + 64 + 64 * 64 + 64] = 0.5
    x = np.zeros((2, 96)); x[0, 5] = 1.5; x[1, 5] = -1.0
    assert oracle.mlp_forward(p, x).tolist() == [4.5, 0.0]


GOLDEN = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "golden.npz"))


@pytest.mark.parametrize("distr", [0, 1, 2, 3])
def test_golden_generator(distr):
    """The library's host generator against the pure-numpy restatement."""
    import topk_amd
    x = topk_amd.generate(300, 8, distr, 7)
    assert np.array_equal(x.view(np.uint32), GOLDEN[f"gen_{distr}"])


def test_golden_topk():
    """The oracle's top-k against the pure-numpy fixtures (reduced C1 and a
    weighted, unordered query with 49 objects tied at the top score)."""
    import topk_amd
    n = 1 << 16
    x = topk_amd.generate(n, 8, 1, 0)
    rid, rs, _ = oracle.fte_topk(x, n, 16, np.ones(8, np.float32), [4, 5, 6, 7])
    assert np.array_equal(rid, GOLDEN["c1_ids"]) and np.array_equal(rs.view(np.uint32), GOLDEN["c1_scores"])
    xc = topk_amd.generate(n, 8, 0, 2)
    rid, rs, _ = oracle.fte_topk(xc, n, 64, GOLDEN["cw_weights"], GOLDEN["cw_attr"])
    assert np.array_equal(rid, GOLDEN["cw_ids"]) and np.array_equal(rs.view(np.uint32), GOLDEN["cw_scores"])
