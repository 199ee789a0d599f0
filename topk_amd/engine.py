"""Python host side over the C ABI (include/topk.h).

:class:`Engine` owns one ``topk_ctx`` (a column store in HBM plus a HIP
stream); the module functions cover data ingest (text and column files).
The reference's algorithm-object interface (TPAc / PTA / File) is mirrored
for C++ callers in include/topk_amd.hpp.
"""
import ctypes
import os

import numpy as np

from ._lib import (MLP_IN, MLP_NPARAMS, TOPK_MAX_DIMS, TOPK_ROWMAJOR, TopkBatchTuning, TopkIngestReport,
                   TopkLearnedConfig, TopkLearnedReport, TopkQuery, TopkStats, TopkStreamReport, check, lib)

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


def _report(rep):
    return {name: getattr(rep, name) for name, _ in TopkIngestReport._fields_}


def csv_count(path):
    """(rows, fields) of a text file: non-blank lines, values on the first one."""
    n = _c.c_uint64()
    d = _c.c_uint32()
    check(lib().topk_csv_count(os.fsencode(path), _c.byref(n), _c.byref(d)), "topk_csv_count")
    return n.value, d.value


def read_csv(path, n=0, d=0, rowmajor=False):
    """Parses a text file on the GPU: returns (table, report), the table
    (d, n) column-major or (n, d) with rowmajor=True."""
    if n == 0 or d == 0:
        cn, cd = csv_count(path)
        n = n or cn
        d = d or cd
    out = np.empty((n, d) if rowmajor else (d, n), dtype=np.float32)
    rep = TopkIngestReport()
    check(lib().topk_csv_read(os.fsencode(path), n, d, TOPK_ROWMAJOR if rowmajor else 0, _ptr(out),
                              _c.byref(rep)), "topk_csv_read")
    return out, _report(rep)


def write_csv(path, data, rowmajor=False):
    """Writes a (d, n) column-major table (or (n, d) with rowmajor=True) as
    text rows of "%f" values."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    n, d = (a.shape if rowmajor else a.shape[::-1])
    check(lib().topk_csv_write(os.fsencode(path), _ptr(a), n, d, TOPK_ROWMAJOR if rowmajor else 0),
          "topk_csv_write")


def write_colfile(path, data, rowmajor=False):
    """Writes a binary column file (4096-byte header, then the columns)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    n, d = (a.shape if rowmajor else a.shape[::-1])
    check(lib().topk_colfile_write(os.fsencode(path), _ptr(a), n, d, TOPK_ROWMAJOR if rowmajor else 0),
          "topk_colfile_write")


def colfile_info(path):
    n = _c.c_uint64()
    d = _c.c_uint32()
    check(lib().topk_colfile_info(os.fsencode(path), _c.byref(n), _c.byref(d)), "topk_colfile_info")
    return n.value, d.value


def _query_arrays(d, weights, attr):
    attr = np.ascontiguousarray(attr, dtype=np.uint8)
    w = np.zeros(max(d, TOPK_MAX_DIMS), dtype=np.float32)
    wa = np.asarray(weights, dtype=np.float32).ravel()
    if wa.size < d:
        raise ValueError("weights are indexed by attribute id and need d entries")
    w[:wa.size] = wa[:w.size]
    return w, attr


def query_array(queries):
    """ctypes array of topk_query from (k, weights_by_attr_id, attr) tuples."""
    nq = len(queries)
    arr = np.zeros(nq, dtype=np.dtype([("k", "<u4"), ("qq", "<u4"), ("attr", "u1", TOPK_MAX_DIMS),
                                       ("weights", "<f4", TOPK_MAX_DIMS)]))
    for i, (k, w, at) in enumerate(queries):
        at = np.asarray(at, dtype=np.uint8).ravel()
        wa = np.asarray(w, dtype=np.float32).ravel()
        if at.size > TOPK_MAX_DIMS:
            raise ValueError("at most 32 terms per query")
        arr[i]["k"] = k
        arr[i]["qq"] = at.size
        arr[i]["attr"][:at.size] = at
        n = min(wa.size, TOPK_MAX_DIMS)
        arr[i]["weights"][:n] = wa[:n]
    qs = (TopkQuery * nq)()
    _c.memmove(qs, arr.ctypes.data, arr.nbytes)
    return qs


def _host_array(count, dtype):
    """An output array in pinned host memory when a GPU is present (the
    library then copies results into it directly), else pageable."""
    try:
        import torch
        if torch.cuda.is_available():
            tdt = {np.uint64: torch.int64, np.float32: torch.float32}[dtype]
            t = torch.empty(count, dtype=tdt, pin_memory=True)
            a = t.numpy().view(dtype)
            return a, t
    except (ImportError, RuntimeError):
        pass
    return np.empty(count, dtype=dtype), None


class PreparedBatch:
    """A query list staged once as topk_query structs plus output arrays."""

    def __init__(self, queries):
        self.queries = queries
        self.nq = len(queries)
        self.qs = query_array(queries)
        self.kmax = max(1, max(int(q[0]) for q in queries))
        self.ids, self._tids = _host_array(self.nq * self.kmax, np.uint64)
        self.scores, self._tsc = _host_array(self.nq * self.kmax, np.float32)
        self.thr, self._tthr = _host_array(self.nq, np.float32)
        # the submit arguments, converted once (a serving loop resubmits them)
        self._args = (self.nq, _c.cast(self.qs, _c.c_void_p), self.kmax, _ptr(self.ids), _ptr(self.scores),
                      _ptr(self.thr))

    def result(self, i, n):
        c = min(int(self.queries[i][0]), n)
        o = i * self.kmax
        return self.ids[o:o + c], self.scores[o:o + c], float(self.thr[i])


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

    @classmethod
    def from_file(cls, path, n=0, d=0):
        """Column store from a text file (one object per line, ',' between
        values; parsed on the GPU) or a binary column file; n / d of 0 take
        the whole file.  The ingest report is kept as ``ingest``."""
        h = _c.c_void_p()
        rep = TopkIngestReport()
        check(lib().topk_ctx_create_file(_c.byref(h), os.fsencode(path), n, d, 0, _c.byref(rep)),
              "topk_ctx_create_file")
        e = cls(_handle=h)
        e.ingest = _report(rep)
        return e

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
        qs = query_array(queries)
        kmax = max(1, max(int(q[0]) for q in queries))
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

    def prepare_batch(self, queries):
        """Staged form of a query list for repeated batch_run calls."""
        return PreparedBatch(queries)

    def batch_run(self, pb):
        """Runs a PreparedBatch; results land in pb.ids / pb.scores / pb.thr."""
        check(lib().topk_batch(self._h, pb.nq, _c.cast(pb.qs, _c.c_void_p), pb.kmax, _ptr(pb.ids), _ptr(pb.scores),
                               _ptr(pb.thr)), "topk_batch")

    def batch_submit(self, pb):
        """Launches a PreparedBatch without waiting; returns its ticket.  The
        results are in pb.ids / pb.scores / pb.thr after batch_wait(ticket)."""
        t = _c.c_uint64()
        check(lib().topk_batch_submit(self._h, *pb._args, _c.byref(t)), "topk_batch_submit")
        return t.value

    def batch_wait(self, ticket):
        check(lib().topk_batch_wait(self._h, ticket), "topk_batch_wait")

    def batch_autotune(self, queries):
        """Times every batch launch shape on these queries and installs the
        fastest; returns the choice as a dict."""
        pb = queries if isinstance(queries, PreparedBatch) else PreparedBatch(queries)
        t = TopkBatchTuning()
        check(lib().topk_batch_autotune(self._h, pb.nq, _c.cast(pb.qs, _c.c_void_p), pb.kmax, _c.byref(t)),
              "topk_batch_autotune")
        return {f: getattr(t, f) for f, _ in TopkBatchTuning._fields_}

    # ---- learned threshold ----
    def learned_train(self, queries, **cfg):
        """Trains and installs the k-th score predictor; returns the report as a dict."""
        c = TopkLearnedConfig()
        check(lib().topk_learned_default_config(_c.byref(c)), "topk_learned_default_config")
        for k, v in cfg.items():
            setattr(c, k, v)
        r = TopkLearnedReport()
        qs = query_array(queries)
        check(lib().topk_learned_train(self._h, len(queries), _c.cast(qs, _c.c_void_p), _c.byref(c), _c.byref(r)),
              "topk_learned_train")
        return {f: getattr(r, f) for f, _ in TopkLearnedReport._fields_}

    def learned_clear(self):
        check(lib().topk_learned_clear(self._h), "topk_learned_clear")

    def learned_get(self):
        p = np.empty(MLP_NPARAMS, np.float32)
        m = _c.c_float()
        check(lib().topk_learned_get(self._h, _ptr(p), _c.byref(m)), "topk_learned_get")
        return p, m.value

    def learned_set(self, params, margin):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != MLP_NPARAMS:
            raise ValueError("10433 parameters expected")
        check(lib().topk_learned_set(self._h, _ptr(p), float(margin)), "topk_learned_set")

    def learned_features(self, queries):
        qs = query_array(queries)
        x = np.empty((len(queries), MLP_IN), np.float32)
        sc = np.empty(len(queries), np.float32)
        check(lib().topk_learned_features(self._h, len(queries), _c.cast(qs, _c.c_void_p), _ptr(x), _ptr(sc)),
              "topk_learned_features")
        return x, sc

    def learned_predict(self, queries):
        qs = query_array(queries)
        pred = np.empty(len(queries), np.float32)
        tau = np.empty(len(queries), np.float32)
        check(lib().topk_learned_predict(self._h, len(queries), _c.cast(qs, _c.c_void_p), _ptr(pred), _ptr(tau)),
              "topk_learned_predict")
        return pred, tau

    def mlp_eval(self, params, x, y=None):
        """GPU forward (and with y, backward): (out, grad, loss)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        xx = np.ascontiguousarray(x, dtype=np.float32)
        nb = xx.shape[0]
        out = np.empty(nb, np.float32)
        grad = np.empty(MLP_NPARAMS, np.float32) if y is not None else None
        loss = _c.c_float()
        yy = np.ascontiguousarray(y, dtype=np.float32) if y is not None else None
        check(lib().topk_mlp_eval(self._h, _ptr(p), _ptr(xx), _ptr(yy) if yy is not None else None, nb, _ptr(out),
                                  _ptr(grad) if grad is not None else None,
                                  _c.byref(loss) if y is not None else None), "topk_mlp_eval")
        return out, grad, (loss.value if y is not None else None)

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


class HostStore:
    """A column store in pinned host memory, streamed through HBM in chunks
    by every query (stores larger than HBM).  ``data`` is the (d, n)
    column-major host array to fill before querying."""

    def __init__(self, n, d, chunk=0):
        self._h = _c.c_void_p()
        check(lib().topk_hstore_create(_c.byref(self._h), n, d, chunk), "topk_hstore_create")
        self.n, self.d = n, d
        ptr = lib().topk_hstore_data(self._h)
        self.data = np.ctypeslib.as_array((_c.c_float * (n * d)).from_address(ptr)).reshape(d, n)

    def close(self):
        if self._h:
            self.data = None
            lib().topk_hstore_destroy(self._h)
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

    def generate(self, distr, seed, threads=0):
        """Fills the host store with the synthetic generator (same bits as
        Engine.generated)."""
        check(lib().topk_generate(lib().topk_hstore_data(self._h), self.n, self.d, self.n, distr, seed, 0, self.n,
                                  threads), "topk_generate")

    def fte(self, k, weights, attr):
        """Exact top-k over the whole host store: (ids, scores, threshold), best first."""
        w, at = _query_arrays(self.d, weights, attr)
        ids = np.empty(k, dtype=np.uint64)
        sc = np.empty(k, dtype=np.float32)
        thr = np.zeros(1, dtype=np.float32)
        check(lib().topk_hstore_fte(self._h, k, at.size, _ptr(w), _ptr(at), _ptr(ids), _ptr(sc), _ptr(thr)),
              "topk_hstore_fte")
        c = min(k, self.n)
        return ids[:c], sc[:c], float(thr[0])

    def batch(self, queries):
        """queries: list of (k, weights_by_attr_id, attr); each chunk crosses
        PCIe once for all of them.  Returns (ids, scores, threshold) per query."""
        nq = len(queries)
        qs = query_array(queries)
        kmax = max(1, max(int(q[0]) for q in queries))
        ids = np.empty(nq * kmax, dtype=np.uint64)
        sc = np.empty(nq * kmax, dtype=np.float32)
        thr = np.empty(nq, dtype=np.float32)
        check(lib().topk_hstore_batch(self._h, nq, _c.cast(qs, _c.c_void_p), kmax, _ptr(ids), _ptr(sc), _ptr(thr)),
              "topk_hstore_batch")
        out = []
        for i, (k, _, _) in enumerate(queries):
            c = min(k, self.n)
            out.append((ids[i * kmax:i * kmax + c].copy(), sc[i * kmax:i * kmax + c].copy(), float(thr[i])))
        return out

    def report(self):
        r = TopkStreamReport()
        check(lib().topk_hstore_report(self._h, _c.byref(r)), "topk_hstore_report")
        return {name: getattr(r, name) for name, _ in TopkStreamReport._fields_}
