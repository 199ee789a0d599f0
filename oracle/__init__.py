"""CPU oracle for the top-k query hot path -- TEST INFRASTRUCTURE ONLY.

Thin numpy/ctypes wrapper over ``liboracle.so`` (built from ``topk_oracle.c``
by ``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker; the product (``topk_amd``) never touches it.

Each function cites the reference code it restates; see ``topk_oracle.c``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c = ctypes


def lib():
    """Load (building first if needed) the oracle shared library."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", _HERE])
    L = ctypes.CDLL(path)
    sig = {
        "or_okey": (_c.c_uint32, [_c.c_float]),
        "or_okey_inv": (_c.c_float, [_c.c_uint32]),
        "or_key": (_c.c_uint64, [_c.c_float, _c.c_uint64]),
        "or_score": (None, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _f32p, _u8p, _f32p]),
        "or_score_div": (None, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _f32p, _u8p, _f32p]),
        "or_fte_topk": (_c.c_uint32, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_uint32,
                                      _f32p, _u8p, _u64p, _f32p, _f32p]),
        "or_fte_topk_div": (_c.c_uint32, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_uint32,
                                          _f32p, _u8p, _u64p, _f32p, _f32p]),
        "or_tpac_threads": (_c.c_uint32, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_uint32,
                                          _f32p, _u8p, _c.c_int, _u64p, _f32p, _f32p]),
        "or_fte_topk_wide": (_c.c_uint32, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _c.c_uint32,
                                           _f32p, _u8p, _c.c_uint64, _c.c_int, _u64p, _f32p, _f32p]),
        "or_mq_threads": (_c.c_int, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _u32p, _u32p,
                                     _f32p, _c.c_uint32, _u8p, _c.c_uint32, _c.c_int,
                                     _u64p, _f32p, _f32p, _c.c_uint32]),
        "or_radix_select_u32": (_c.c_uint32, [_u32p, _c.c_uint64, _c.c_uint64]),
        "or_radix_select_u64": (_c.c_uint64, [_u64p, _c.c_uint64, _c.c_uint64]),
        "or_normalize": (None, [_f32p, _c.c_uint64, _c.c_uint32, _c.c_uint64]),
        "or_reorder_attr": (None, [_f32p, _c.c_uint64, _c.c_uint32, _c.c_uint64]),
        "or_check_order": (_c.c_uint64, [_f32p, _c.c_uint64, _c.c_uint32, _c.c_uint64]),
        "or_bta_blocks": (None, [_f32p, _c.c_uint64, _c.c_uint64, _c.c_uint32, _f32p, _u8p,
                                 _c.c_uint32, _c.c_uint32, _u64p]),
        "or_gpa_step": (_c.c_uint64, [_u32p, _f32p, _c.c_uint64, _f32p, _f32p, _c.c_float, _c.c_float]),
        "or_pta_bound": (_c.c_float, [_f32p, _f32p, _c.c_uint32, _c.c_uint64, _c.c_uint32, _f32p, _u8p]),
        "or_pta_topk": (_c.c_uint32, [_f32p, _u32p, _c.c_uint64, _c.c_uint32, _c.c_uint32, _u32p, _u32p,
                                      _f32p, _f32p, _c.c_uint32, _c.c_uint32, _f32p, _u8p,
                                      _u64p, _f32p, _f32p, _c.POINTER(_c.c_uint64)]),
        "or_pta_threads": (_c.c_uint32, [_f32p, _u32p, _c.c_uint64, _c.c_uint32, _c.c_uint32, _u32p, _u32p,
                                         _f32p, _f32p, _c.c_uint32, _c.c_uint32, _f32p, _u8p, _c.c_int,
                                         _u64p, _f32p, _f32p, _c.POINTER(_c.c_uint64)]),
        "or_mlp_nparams": (_c.c_uint32, []),
        "or_mlp_forward": (None, [_f64p, _f64p, _c.c_uint32, _f64p]),
        "or_mlp_backward": (_c.c_double, [_f64p, _f64p, _f64p, _c.c_uint32, _f64p]),
        "or_csv_count": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint32)]),
        "or_csv_read": (_c.c_int64, [_c.c_char_p, _c.c_uint64, _c.c_uint32, _c.c_int, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _cols(data):
    """(d, ld) float32 column-major view -> (array, d, ld)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError("data must be (d, n) column-major")
    return a, a.shape[0], a.shape[1]


def _query(d, weights, attr):
    w = np.ascontiguousarray(weights, dtype=np.float32)
    at = np.ascontiguousarray(attr, dtype=np.uint8)
    if w.size < d:
        raise ValueError("weights are indexed by attribute id and need >= d entries")
    return w, at


def okey(s):
    return lib().or_okey(float(s))


def key(s, i):
    return lib().or_key(float(s), int(i))


def score(data, n, weights, attr, divide=False):
    """Per-object scores, cpu/TPAc.h:129-142 order; divide=True divides the
    running score by qq after every term (LD == 2, cpu/TPAc.h:143-146)."""
    a, d, ld = _cols(data)
    w, at = _query(d, weights, attr)
    out = np.empty(n, dtype=np.float32)
    (lib().or_score_div if divide else lib().or_score)(a.ravel(), n, ld, at.size, w, at, out)
    return out


def fte_topk(data, n, k, weights, attr, divide=False):
    """Scalar FTE top-k (cpu/TPAc.h:35-207 restated).  Returns (ids, scores, thr).
    divide=True: real-data scores (score(..., divide=True))."""
    a, d, ld = _cols(data)
    w, at = _query(d, weights, attr)
    ids = np.zeros(k, dtype=np.uint64)
    sc = np.zeros(k, dtype=np.float32)
    thr = np.zeros(1, dtype=np.float32)
    fn = lib().or_fte_topk_div if divide else lib().or_fte_topk
    c = fn(a.ravel(), n, ld, k, at.size, w, at, ids, sc, thr)
    return ids[:c], sc[:c], float(thr[0])


def tpac_threads(data, n, k, weights, attr, threads):
    """SIMD + threads FTE baseline (cpu/TPAc.h:256-358 restated)."""
    a, d, ld = _cols(data)
    w, at = _query(d, weights, attr)
    ids = np.zeros(k, dtype=np.uint64)
    sc = np.zeros(k, dtype=np.float32)
    thr = np.zeros(1, dtype=np.float32)
    c = lib().or_tpac_threads(a.ravel(), n, ld, k, at.size, w, at, threads, ids, sc, thr)
    return ids[:c], sc[:c], float(thr[0])


def fte_topk_wide(data, n, k, weights, attr, threads=1, id_base=0):
    """FTE top-k with 64-bit ids (object i has id id_base + i; ties by the
    lower 64-bit id), over threads.  Returns (ids, scores, thr)."""
    a, d, ld = _cols(data)
    w, at = _query(d, weights, attr)
    ids = np.zeros(k, dtype=np.uint64)
    sc = np.zeros(k, dtype=np.float32)
    thr = np.zeros(1, dtype=np.float32)
    c = lib().or_fte_topk_wide(a.ravel(), n, ld, k, at.size, w, at, id_base, threads, ids, sc, thr)
    return ids[:c], sc[:c], float(thr[0])


def mq_threads(data, n, ks, qqs, weights, attrs, threads):
    """One full scan per query, queries over threads (cpu/TPAc.h:209-254).

    weights: (nq, d) by attribute id; attrs: (nq, amax) with the first qq used.
    Returns lists of (ids, scores) per query (scores desc, id asc)."""
    a, d, ld = _cols(data)
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    qqs = np.ascontiguousarray(qqs, dtype=np.uint32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    at = np.ascontiguousarray(attrs, dtype=np.uint8)
    nq = ks.size
    kmax = int(ks.max())
    ids = np.zeros(nq * kmax, dtype=np.uint64)
    sc = np.zeros(nq * kmax, dtype=np.float32)
    thr = np.zeros(nq, dtype=np.float32)
    lib().or_mq_threads(a.ravel(), n, ld, nq, ks, qqs, w.ravel(), w.shape[1], at.ravel(),
                        at.shape[1], threads, ids, sc, thr, kmax)
    out = []
    for q in range(nq):
        c = min(int(ks[q]), n)
        out.append((ids[q * kmax:q * kmax + c].copy(), sc[q * kmax:q * kmax + c].copy()))
    return out


def radix_select_u32(keys, r):
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    return lib().or_radix_select_u32(k, k.size, r)


def radix_select_u64(keys, r):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    return lib().or_radix_select_u64(k, k.size, r)


def normalize(data, n):
    a, d, ld = _cols(data)
    a = a.copy()
    lib().or_normalize(a.ravel(), n, d, ld)
    return a


def reorder_attr(data, n):
    a, d, ld = _cols(data)
    a = a.copy()
    lib().or_reorder_attr(a.ravel(), n, d, ld)
    return a


def check_order(data, n):
    a, d, ld = _cols(data)
    return lib().or_check_order(a.ravel(), n, d, ld)


def bta_blocks(data, n, weights, attr, block, k):
    a, d, ld = _cols(data)
    w, at = _query(d, weights, attr)
    nb = (n + block - 1) // block
    out = np.zeros(nb * k, dtype=np.uint64)
    lib().or_bta_blocks(a.ravel(), n, ld, at.size, w, at, block, k, out)
    return out.reshape(nb, k)


def gpa_step(ids, score, colC, colN, thr, suffix_len):
    ids = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    sc = np.ascontiguousarray(score, dtype=np.float32).copy()
    cC = np.ascontiguousarray(colC, dtype=np.float32)
    cN = np.ascontiguousarray(colN if colN is not None else colC, dtype=np.float32)
    fn = lib().or_gpa_step
    if colN is None:
        fn.argtypes = [_u32p, _f32p, _c.c_uint64, _f32p, _c.c_void_p, _c.c_float, _c.c_float]
        c = fn(ids, sc, ids.size, cC, None, thr, suffix_len)
        fn.argtypes = [_u32p, _f32p, _c.c_uint64, _f32p, _f32p, _c.c_float, _c.c_float]
    else:
        c = fn(ids, sc, ids.size, cC, cN, thr, suffix_len)
    return ids[:c], sc[:c]


def pta_topk(pdata, pid, poff, psz, pmax, pmin, k, weights, attr, threads=0):
    """PTA over an explicit partition layout (cpu/PTA.h:357-422, 569-671).

    pdata: (d, ld) partition-order columns; partition p is positions
    [poff[p], poff[p]+psz[p]).  Returns (ids, scores, thr, scored_partitions)."""
    a, d, ld = _cols(pdata)
    w, at = _query(d, weights, attr)
    pid = np.ascontiguousarray(pid, dtype=np.uint32)
    poff = np.ascontiguousarray(poff, dtype=np.uint32)
    psz = np.ascontiguousarray(psz, dtype=np.uint32)
    nparts = psz.size
    pmax = np.ascontiguousarray(pmax, dtype=np.float32).ravel()
    pmin = np.ascontiguousarray(pmin, dtype=np.float32).ravel()
    ids = np.zeros(k, dtype=np.uint64)
    sc = np.zeros(k, dtype=np.float32)
    thr = np.zeros(1, dtype=np.float32)
    scored = _c.c_uint64(0)
    if threads:
        c = lib().or_pta_threads(a.ravel(), pid, ld, d, nparts, poff, psz, pmax, pmin, k, at.size, w, at,
                                 threads, ids, sc, thr, _c.byref(scored))
    else:
        c = lib().or_pta_topk(a.ravel(), pid, ld, d, nparts, poff, psz, pmax, pmin, k, at.size, w, at,
                              ids, sc, thr, _c.byref(scored))
    return ids[:c], sc[:c], float(thr[0]), int(scored.value)


def pta_bound(pmax, pmin, d, p, weights, attr):
    w, at = _query(d, weights, attr)
    return lib().or_pta_bound(np.ascontiguousarray(pmax, dtype=np.float32).ravel(),
                              np.ascontiguousarray(pmin, dtype=np.float32).ravel(),
                              d, p, at.size, w, at)


def mlp_nparams():
    return lib().or_mlp_nparams()


def mlp_forward(params, x):
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(xx.shape[0], dtype=np.float64)
    lib().or_mlp_forward(p, xx.ravel(), xx.shape[0], out)
    return out


def mlp_backward(params, x, y):
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    yy = np.ascontiguousarray(y, dtype=np.float64)
    g = np.zeros(p.size, dtype=np.float64)
    loss = lib().or_mlp_backward(p, xx.ravel(), yy, xx.shape[0], g)
    return loss, g


def csv_count(path):
    """(n, d) as the reference's File::count computes them (input/File.h:107-128)."""
    n = _c.c_uint64()
    d = _c.c_uint32()
    if lib().or_csv_count(os.fsencode(path), _c.byref(n), _c.byref(d)) != 0:
        raise OSError(path)
    return n.value, d.value


def csv_read(path, n, d, colmajor=True):
    """The reference's fscanf row loop (input/File.h:233-295): returns the
    table (d x n column-major as a 2-D array, or n x d) and the rows read;
    raises ValueError for a row with fewer than d values."""
    out = np.zeros((d, n) if colmajor else (n, d), np.float32)
    r = lib().or_csv_read(os.fsencode(path), n, d, 1 if colmajor else 0, out)
    if r == -2:
        raise OSError(path)
    if r < 0:
        raise ValueError("row with fewer than %d values" % d)
    return out, int(r)
