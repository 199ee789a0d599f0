"""Python host side over the C ABI (include/topk.h).

:class:`Engine` owns one ``topk_ctx`` (a column store in HBM plus a HIP
stream).  :mod:`topk_amd.compat` mirrors the reference's algorithm-object
interface (TPAc / PTA / BTA / GPA) on top of it.
"""
import ctypes

import numpy as np

from ._lib import (TOPK_MAX_DIMS, TOPK_ROWMAJOR, TopkQuery, TopkStats, check, lib)

_c = ctypes


def _ptr(a):
    return a.ctypes.data_as(_c.c_void_p)


def generate(n, d, distr, seed, i0=0, count=None, threads=0):
    """Objects [i0, i0+count) of the synthetic n x d dataset as a (d, count)
    float32 column-major array (same bits as the device generator)."""
    if count is None:
        count = n - i0
    out = np.empty((d, count), dtype=np.float32)
    check(lib().topk_generate(_ptr(out), n, d, count, distr, seed, i0, count, threads), "topk_generate")
    return out


def _query_arrays(d, weights, attr):
    attr = np.ascontiguousarray(attr, dtype=np.uint8)
    w = np.zeros(max(d, TOPK_MAX_DIMS), dtype=np.float32)
    wa = np.asarray(weights, dtype=np.float32).ravel()
    if wa.size < d:
        raise ValueError("weights are indexed by attribute id and need d entries")
    w[:wa.size] = wa[:w.size]
    return w, attr


class Engine:
    """One top-k context: column store + stream on the current GPU."""

    def __init__(self, data=None, *, rowmajor=False, _handle=None):
        self._h = _c.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            a = np.ascontiguousarray(data, dtype=np.float32)
            if a.ndim != 2:
                raise ValueError("data must be 2-D: (d, n) column-major or (n, d) with rowmajor=True")
            if rowmajor:
                n, d = a.shape
            else:
                d, n = a.shape
            check(lib().topk_ctx_create(_c.byref(self._h), _ptr(a), n, d, TOPK_ROWMAJOR if rowmajor else 0),
                  "topk_ctx_create")
        self.n = lib().topk_ctx_n(self._h)
        self.d = lib().topk_ctx_d(self._h)

    @classmethod
    def generated(cls, n, d, distr, seed):
        h = _c.c_void_p()
        check(lib().topk_ctx_create_generated(_c.byref(h), n, d, distr, seed), "topk_ctx_create_generated")
        return cls(_handle=h)

    def close(self):
        if self._h:
            lib().topk_ctx_destroy(self._h)
            self._h = _c.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # ---- store / init ----
    def read_columns(self):
        out = np.empty((self.d, self.n), dtype=np.float32)
        check(lib().topk_ctx_read_columns(self._h, _ptr(out)), "topk_ctx_read_columns")
        return out

    def init(self, flags):
        check(lib().topk_init(self._h, flags), "topk_init")

    def set_partition_size(self, objects):
        check(lib().topk_set_partition_size(self._h, objects), "topk_set_partition_size")

    def reorder_attr(self):
        check(lib().topk_reorder_attr(self._h), "topk_reorder_attr")

    def pta_layout(self):
        """The PTA partition layout as numpy arrays:
        (pid[npos], poff[nparts+1], psz[nparts], pmax[nparts,d], pmin[nparts,d])."""
        np_ = _c.c_uint32()
        npos = _c.c_uint64()
        check(lib().topk_pta_layout(self._h, _c.byref(np_), _c.byref(npos)), "topk_pta_layout")
        P, N = np_.value, npos.value
        pid = np.empty(N, np.uint32)
        poff = np.empty(P + 1, np.uint32)
        psz = np.empty(P, np.uint32)
        pmax = np.empty((P, self.d), np.float32)
        pmin = np.empty((P, self.d), np.float32)
        check(lib().topk_pta_read_layout(self._h, _ptr(pid), _ptr(poff), _ptr(psz), _ptr(pmax), _ptr(pmin)),
              "topk_pta_read_layout")
        return pid, poff, psz, pmax, pmin

    # ---- queries ----
    def _single(self, fn, name, k, weights, attr):
        w, at = _query_arrays(self.d, weights, attr)
        ids = np.empty(k, dtype=np.uint64)
        sc = np.empty(k, dtype=np.float32)
        thr = np.zeros(1, dtype=np.float32)
        check(fn(self._h, k, at.size, _ptr(w), _ptr(at), _ptr(ids), _ptr(sc), _ptr(thr)), name)
        c = min(k, self.n)
        return ids[:c], sc[:c], float(thr[0])

    def fte(self, k, weights, attr):
        """Plain full-table top-k: (ids, scores, threshold), best first."""
        return self._single(lib().topk_fte, "topk_fte", k, weights, attr)

    def pta(self, k, weights, attr):
        return self._single(lib().topk_pta, "topk_pta", k, weights, attr)

    def bta(self, k, weights, attr):
        return self._single(lib().topk_bta, "topk_bta", k, weights, attr)

    def gpa(self, k, weights, attr):
        return self._single(lib().topk_gpa, "topk_gpa", k, weights, attr)

    def batch(self, queries):
        """queries: list of (k, weights_by_attr_id, attr).  Returns a list of
        (ids, scores, threshold) per query."""
        nq = len(queries)
        qs = (TopkQuery * nq)()
        kmax = 1
        for i, (k, w, at) in enumerate(queries):
            at = np.asarray(at, dtype=np.uint8)
            wa = np.asarray(w, dtype=np.float32).ravel()
            qs[i].k = k
            qs[i].qq = at.size
            for j, a in enumerate(at):
                qs[i].attr[j] = int(a)
            for j in range(min(wa.size, TOPK_MAX_DIMS)):
                qs[i].weights[j] = float(wa[j])
            kmax = max(kmax, k)
        ids = np.empty(nq * kmax, dtype=np.uint64)
        sc = np.empty(nq * kmax, dtype=np.float32)
        thr = np.empty(nq, dtype=np.float32)
        check(lib().topk_batch(self._h, nq, _c.cast(qs, _c.c_void_p), kmax, _ptr(ids), _ptr(sc), _ptr(thr)),
              "topk_batch")
        out = []
        for i, (k, _, _) in enumerate(queries):
            c = min(k, self.n)
            out.append((ids[i * kmax:i * kmax + c].copy(), sc[i * kmax:i * kmax + c].copy(), float(thr[i])))
        return out

    # ---- device-pointer API (outputs in HBM; callers pass device addresses) ----
    def fte_async(self, k, weights, attr, d_ids, d_scores):
        w, at = _query_arrays(self.d, weights, attr)
        check(lib().topk_fte_async(self._h, k, at.size, _ptr(w), _ptr(at), _c.c_void_p(d_ids),
                                   _c.c_void_p(d_scores)), "topk_fte_async")

    def sync(self):
        check(lib().topk_sync(self._h), "topk_sync")

    def stream(self):
        return lib().topk_stream(self._h)

    def set_tuning(self, name, value):
        check(lib().topk_set_tuning(self._h, name.encode(), int(value)), "topk_set_tuning")

    # ---- stats / timing ----
    def stats(self):
        s = TopkStats()
        check(lib().topk_get_stats(self._h, _c.byref(s)), "topk_get_stats")
        return {f: getattr(s, f) for f, _ in TopkStats._fields_}

    def timing(self, on=True):
        check(lib().topk_timing_enable(self._h, 1 if on else 0), "topk_timing_enable")

    def timing_get(self, kernel):
        ms = _c.c_double()
        cnt = _c.c_uint64()
        check(lib().topk_timing_get(self._h, kernel.encode(), _c.byref(ms), _c.byref(cnt)), "topk_timing_get")
        return ms.value, cnt.value

    def timing_reset(self):
        check(lib().topk_timing_reset(self._h), "topk_timing_reset")

    def radix_select_u32(self, d_keys, n, r):
        out = _c.c_uint32()
        check(lib().topk_radix_select_u32(self._h, _c.c_void_p(d_keys), n, r, _c.byref(out)),
              "topk_radix_select_u32")
        return out.value
