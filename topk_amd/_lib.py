"""ctypes binding of ``libtopk_amd.so`` (the C ABI declared in include/topk.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C topk_amd/csrc``).  There is no fallback: if the shared library is
missing, :func:`lib` raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtopk_amd.so")

TOPK_MAX_DIMS = 32
TOPK_MAX_K = 1024

TOPK_OK = 0
TOPK_ERR_INVALID = 1
TOPK_ERR_NOMEM = 2
TOPK_ERR_HIP = 3
TOPK_ERR_UNSUPPORTED = 4
TOPK_ERR_NONFINITE = 5
TOPK_ERR_STATE = 6
TOPK_ERR_NO_DEVICE = 7
TOPK_ERR_OVERFLOW = 8
TOPK_ERR_FORMAT = 9
TOPK_ERR_IO = 10

TOPK_ROWMAJOR = 0x1
TOPK_INIT_NORMALIZE = 0x1
TOPK_INIT_PTA = 0x2

CORRELATED, INDEPENDENT, ANTICORRELATED, MIXED = 0, 1, 2, 3

_c = ctypes
_vp = _c.c_void_p


class TopkStats(ctypes.Structure):
    _fields_ = [("tuple_count", _c.c_uint64), ("candidates", _c.c_uint64),
                ("partitions_scored", _c.c_uint64), ("partitions_total", _c.c_uint64),
                ("fallback", _c.c_uint32), ("queries", _c.c_uint32),
                ("ms_last", _c.c_double), ("tt_init", _c.c_double),
                ("retried", _c.c_uint32), ("items", _c.c_uint32), ("fte_fallback", _c.c_uint32),
                ("terms", _c.c_uint64)]


class TopkQuery(ctypes.Structure):
    _fields_ = [("k", _c.c_uint32), ("qq", _c.c_uint32),
                ("attr", _c.c_uint8 * TOPK_MAX_DIMS), ("weights", _c.c_float * TOPK_MAX_DIMS)]


class TopkBatchTuning(ctypes.Structure):
    _fields_ = [("vpt", _c.c_uint32), ("bs", _c.c_uint32), ("qpi", _c.c_uint32), ("digit", _c.c_uint32),
                ("grid", _c.c_uint32), ("item", _c.c_uint32), ("variants", _c.c_uint32), ("ms_best", _c.c_double),
                ("ms_default", _c.c_double), ("fused", _c.c_uint32)]


class TopkLearnedConfig(ctypes.Structure):
    _fields_ = [("epochs", _c.c_uint32), ("batch", _c.c_uint32), ("lr", _c.c_float),
                ("target_fallback", _c.c_float), ("calib_fraction", _c.c_float), ("seed", _c.c_uint64)]


class TopkLearnedReport(ctypes.Structure):
    _fields_ = [("n_train", _c.c_uint32), ("n_calib", _c.c_uint32), ("train_mse", _c.c_float),
                ("calib_mse", _c.c_float), ("margin", _c.c_float), ("calib_fallback", _c.c_float),
                ("ms_label", _c.c_double), ("ms_train", _c.c_double), ("saturated", _c.c_uint32),
                ("steps", _c.c_uint32), ("flops", _c.c_double), ("ms_steps", _c.c_double)]


class TopkIngestReport(ctypes.Structure):
    _fields_ = [("format", _c.c_uint32), ("fields", _c.c_uint32), ("bytes", _c.c_uint64),
                ("lines", _c.c_uint64), ("rows", _c.c_uint64), ("slow_lines", _c.c_uint64),
                ("ms_read", _c.c_double), ("ms_parse", _c.c_double), ("ms_total", _c.c_double)]


class TopkStreamReport(ctypes.Structure):
    _fields_ = [("chunks", _c.c_uint64), ("bytes", _c.c_uint64), ("queries", _c.c_uint64),
                ("ms_total", _c.c_double)]


MLP_NPARAMS = 10433
MLP_IN = 96

# name -> (restype, argtypes); mirrors include/topk.h one to one
SIGNATURES = {
    "topk_abi_version": (_c.c_int, []),
    "topk_strerror": (_c.c_char_p, [_c.c_int]),
    "topk_ctx_create": (_c.c_int, [_c.POINTER(_vp), _vp, _c.c_uint64, _c.c_uint32, _c.c_uint32]),
    "topk_ctx_create_generated": (_c.c_int, [_c.POINTER(_vp), _c.c_uint64, _c.c_uint32, _c.c_int,
                                             _c.c_uint64]),
    "topk_ctx_destroy": (_c.c_int, [_vp]),
    "topk_csv_count": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint32)]),
    "topk_csv_read": (_c.c_int, [_c.c_char_p, _c.c_uint64, _c.c_uint32, _c.c_uint32, _vp,
                                 _c.POINTER(TopkIngestReport)]),
    "topk_csv_write": (_c.c_int, [_c.c_char_p, _vp, _c.c_uint64, _c.c_uint32, _c.c_uint32]),
    "topk_colfile_write": (_c.c_int, [_c.c_char_p, _vp, _c.c_uint64, _c.c_uint32, _c.c_uint32]),
    "topk_colfile_info": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint32)]),
    "topk_hstore_create": (_c.c_int, [_c.POINTER(_vp), _c.c_uint64, _c.c_uint32, _c.c_uint64]),
    "topk_hstore_data": (_vp, [_vp]),
    "topk_hstore_n": (_c.c_uint64, [_vp]),
    "topk_hstore_d": (_c.c_uint32, [_vp]),
    "topk_hstore_destroy": (_c.c_int, [_vp]),
    "topk_hstore_fte": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp, _vp]),
    "topk_hstore_batch": (_c.c_int, [_vp, _c.c_uint32, _vp, _c.c_uint32, _vp, _vp, _vp]),
    "topk_hstore_report": (_c.c_int, [_vp, _c.POINTER(TopkStreamReport)]),
    "topk_ctx_create_file": (_c.c_int, [_c.POINTER(_vp), _c.c_char_p, _c.c_uint64, _c.c_uint32, _c.c_uint32,
                                        _c.POINTER(TopkIngestReport)]),
    "topk_ctx_n": (_c.c_uint64, [_vp]),
    "topk_ctx_d": (_c.c_uint32, [_vp]),
    "topk_ctx_read_columns": (_c.c_int, [_vp, _vp]),
    "topk_init": (_c.c_int, [_vp, _c.c_uint32]),
    "topk_set_partition_size": (_c.c_int, [_vp, _c.c_uint32]),
    "topk_reorder_attr": (_c.c_int, [_vp]),
    "topk_pta_layout": (_c.c_int, [_vp, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint64)]),
    "topk_pta_read_layout": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "topk_fte": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp, _vp]),
    "topk_pta": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp, _vp]),
    "topk_bta": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp, _vp]),
    "topk_gpa": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp, _vp]),
    "topk_batch": (_c.c_int, [_vp, _c.c_uint32, _vp, _c.c_uint32, _vp, _vp, _vp]),
    "topk_batch_submit": (_c.c_int, [_vp, _c.c_uint32, _vp, _c.c_uint32, _vp, _vp, _vp, _c.POINTER(_c.c_uint64)]),
    "topk_batch_wait": (_c.c_int, [_vp, _c.c_uint64]),
    "topk_fte_async": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint8, _vp, _vp, _vp, _vp]),
    "topk_fte_noalloc": (_c.c_int, [_vp, _c.c_uint64, _c.c_uint32, _c.c_uint64, _c.c_uint32, _c.c_uint8, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _vp]),
    "topk_sync": (_c.c_int, [_vp]),
    "topk_stream": (_vp, [_vp]),
    "topk_set_tuning": (_c.c_int, [_vp, _c.c_char_p, _c.c_int64]),
    "topk_get_stats": (_c.c_int, [_vp, _c.POINTER(TopkStats)]),
    "topk_timing_enable": (_c.c_int, [_vp, _c.c_int]),
    "topk_timing_get": (_c.c_int, [_vp, _c.c_char_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_uint64)]),
    "topk_timing_reset": (_c.c_int, [_vp]),
    "topk_generate": (_c.c_int, [_vp, _c.c_uint64, _c.c_uint32, _c.c_uint64, _c.c_int, _c.c_uint64,
                                 _c.c_uint64, _c.c_uint64, _c.c_int]),
    "topk_radix_select_u32": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.c_uint64, _c.POINTER(_c.c_uint32)]),
    "topk_batch_autotune": (_c.c_int, [_vp, _c.c_uint32, _vp, _c.c_uint32, _c.POINTER(TopkBatchTuning)]),
    "topk_learned_default_config": (_c.c_int, [_c.POINTER(TopkLearnedConfig)]),
    "topk_learned_train": (_c.c_int, [_vp, _c.c_uint32, _vp, _c.POINTER(TopkLearnedConfig),
                                      _c.POINTER(TopkLearnedReport)]),
    "topk_learned_clear": (_c.c_int, [_vp]),
    "topk_learned_get": (_c.c_int, [_vp, _vp, _c.POINTER(_c.c_float)]),
    "topk_learned_set": (_c.c_int, [_vp, _vp, _c.c_float]),
    "topk_learned_features": (_c.c_int, [_vp, _c.c_uint32, _vp, _vp, _vp]),
    "topk_learned_predict": (_c.c_int, [_vp, _c.c_uint32, _vp, _vp, _vp]),
    "topk_mlp_eval": (_c.c_int, [_vp, _vp, _vp, _vp, _c.c_uint32, _vp, _vp, _c.POINTER(_c.c_float)]),
}

_LIB = None


class TopkError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().topk_strerror(status).decode()
        super().__init__(f"{what}: topk status {status} ({msg})" if what else f"topk status {status} ({msg})")


def lib():
    """Load libtopk_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or `make -C topk_amd/csrc` (there is no CPU fallback)")
    # A process that also uses PyTorch must share its HIP runtime: load
    # torch's first so the library binds to the same libamdhip64 (two
    # runtimes in one process leave the second without a usable device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(status, what=""):
    if status != TOPK_OK:
        raise TopkError(status, what)
    return status
