"""topk_fte_noalloc (the allocation-free FTE over a caller's device store,
scratch in a per-GPU device array) against topk_fte on the same data: three
calls in a row sharing that scratch -- a sample-seeded store, then a smaller
one (exact-prefix seed), then a store whose sample guess is planted too high
so the device-side exact re-run answers it."""
import ctypes

import numpy as np
import pytest

import oracle
import topk_amd
from topk_amd import _lib

pytestmark = pytest.mark.gpu


def _device_store(x, ld):
    """Column m of x at m*ld of a zero-padded device buffer of d*ld + 8192 floats."""
    import torch
    d, n = x.shape
    buf = torch.zeros(d * ld + 8192, dtype=torch.float32, device="cuda")
    host = np.zeros(d * ld + 8192, np.float32)
    for m in range(d):
        host[m * ld:m * ld + n] = x[m]
    buf.copy_(torch.from_numpy(host))
    return buf


def _noalloc(buf, n, d, ld, k, w, attr):
    import torch
    L = _lib.lib()
    ids = torch.full((k,), -7, dtype=torch.int64, device="cuda")
    sc = torch.full((k,), 7.0, dtype=torch.float32, device="cuda")
    thr = torch.full((1,), 7.0, dtype=torch.float32, device="cuda")
    st = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    wv = np.ascontiguousarray(w, np.float32)
    av = np.ascontiguousarray(attr, np.uint8)
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.topk_fte_noalloc(ctypes.c_void_p(buf.data_ptr()), n, d, ld, k, av.size, wv.ctypes.data_as(ctypes.c_void_p),
                            av.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ids.data_ptr()),
                            ctypes.c_void_p(sc.data_ptr()), ctypes.c_void_p(thr.data_ptr()),
                            ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(stream))
    assert rc == _lib.TOPK_OK, rc
    torch.cuda.synchronize()
    return (ids.cpu().numpy().view(np.uint64), sc.cpu().numpy(), np.float32(thr.cpu().numpy()[0]),
            int(st.cpu().numpy()[0]))


def _check_against_fte(x, k, w, attr, ld):
    d, n = x.shape
    buf = _device_store(x, ld)
    ids, sc, thr, status = _noalloc(buf, n, d, ld, k, w, attr)
    assert status == _lib.TOPK_OK
    with topk_amd.Engine(x) as eng:
        rid, rsc, rthr = eng.fte(k, w, attr)
    assert np.array_equal(ids, rid), (n, k, attr)
    assert np.array_equal(sc.view(np.uint32), rsc.view(np.uint32)), (n, k, attr)
    assert thr.tobytes() == np.float32(rthr).tobytes()
    return ids, sc, thr


def test_noalloc_matches_fte_on_consecutive_calls(gpu):
    d = 8
    rng = np.random.default_rng(11)
    # 1: sample-seeded store (n > 2^19), unpadded stride ld = n rounded to 4
    n1 = 3_000_017
    x1 = topk_amd.generate(n1, d, topk_amd.INDEPENDENT, 0)
    w1 = rng.uniform(0.1, 1.0, d).astype(np.float32)
    ids1, sc1, _ = _check_against_fte(x1, 64, w1, [4, 5, 6, 7], (n1 + 3) // 4 * 4)
    rid, rsc, _ = oracle.fte_topk(x1, n1, 64, w1, [4, 5, 6, 7])
    assert np.array_equal(ids1, rid) and np.array_equal(sc1, rsc)
    # 2: a smaller store (exact-prefix seed), another k and attribute order
    n2 = 100_003
    x2 = topk_amd.generate(n2, d, topk_amd.ANTICORRELATED, 2)
    w2 = rng.uniform(0.1, 1.0, d).astype(np.float32)
    _check_against_fte(x2, 100, w2, [7, 2, 5], n2 + 1)
    # 3: the sample pass visits 128 chunks of 1024 objects at a stride of
    # (n / 128) rounded down to 1024; one planted all-ones object per chunk
    # outranks everything, so its guess admits ks (= 8k*2^17/n) of the 128
    # planted objects, fewer than k: the device re-runs the query exactly
    n3 = 2_000_000
    x3 = topk_amd.generate(n3, d, topk_amd.INDEPENDENT, 5)
    stride = ((n3 // 128) // 1024) * 1024
    for b in range(128):
        x3[:, b * stride + 5] = 1.0
    k3 = 200
    assert (8 * k3 * 131072 + n3 - 1) // n3 < 128 < k3 + 1000
    ids3, sc3, _ = _check_against_fte(x3, k3, np.ones(d, np.float32), list(range(d)), n3)
    assert set(int(i) for i in ids3[:128]) == {b * stride + 5 for b in range(128)}
    # the first store again, now that the other calls left their state behind
    ids1b, sc1b, _ = _check_against_fte(x1, 64, w1, [4, 5, 6, 7], n1 + 7)
    assert np.array_equal(ids1b, ids1) and np.array_equal(sc1b, sc1)


def test_noalloc_rejects_bad_sizes(gpu):
    import torch
    L = _lib.lib()
    buf = torch.zeros(4 * 1024 + 8192, dtype=torch.float32, device="cuda")
    out = torch.zeros(2048, dtype=torch.int64, device="cuda")
    w = np.ones(4, np.float32)
    a = np.arange(4, dtype=np.uint8)
    vp = ctypes.c_void_p
    args = lambda n, d, ld, k, qq: (vp(buf.data_ptr()), n, d, ld, k, qq, w.ctypes.data_as(vp), a.ctypes.data_as(vp),
                                    vp(out.data_ptr()), vp(out.data_ptr()), None, vp(out.data_ptr()), None)
    assert L.topk_fte_noalloc(*args(1000, 4, 1024, _lib.TOPK_MAX_K + 1, 4)) == _lib.TOPK_ERR_UNSUPPORTED
    assert L.topk_fte_noalloc(*args(1000, 4, 999, 16, 4)) == _lib.TOPK_ERR_INVALID      # ld < n
    assert L.topk_fte_noalloc(*args(1000, 4, 1022, 16, 4)) == _lib.TOPK_ERR_INVALID     # ld not a multiple of 4
    assert L.topk_fte_noalloc(*args(1000, 4, 1024, 16, 5)) == _lib.TOPK_ERR_INVALID     # qq > d
    assert L.topk_fte_noalloc(*args(0, 4, 1024, 16, 4)) == _lib.TOPK_ERR_INVALID       # empty store
