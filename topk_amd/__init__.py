"""MI355X (gfx950) top-k query engine.

Scores each of N objects as a weighted sum over the queried attributes of a
column-major FP32 store and returns the k highest, ranked by (score desc,
id asc), bit-exact with the reference's CPU TPAc/PTA arithmetic
(avionicscode/TopK cpu/TPAc.h, cpu/PTA.h).  All compute runs in the HIP
library ``libtopk_amd.so`` behind the C ABI of ``include/topk.h``.
"""
from ._lib import (ANTICORRELATED, CORRELATED, INDEPENDENT, MIXED, TOPK_INIT_NORMALIZE, TOPK_INIT_PTA,
                   TOPK_MAX_DIMS, TOPK_MAX_K, TopkError, lib)
from .engine import Engine, HostStore, colfile_info, csv_count, generate, read_csv, write_colfile, write_csv

__all__ = ["Engine", "HostStore", "generate", "lib", "csv_count", "read_csv", "write_csv", "write_colfile", "colfile_info", "TopkError", "CORRELATED", "INDEPENDENT", "ANTICORRELATED", "MIXED",
           "TOPK_INIT_NORMALIZE", "TOPK_INIT_PTA", "TOPK_MAX_DIMS", "TOPK_MAX_K"]
