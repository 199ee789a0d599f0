/*
 * topk.h -- C ABI of the MI355X (gfx950) top-k query engine.
 *
 * Drop-in boundary for the reference's query contract
 * (avionicscode/TopK validation/bench_cpu.h:155-237 bench_tpac and
 * :323-402 bench_pta): an object owning an N x d FP32 column-major
 * attribute store, a separately timed init(), and per-query
 *   findTopK(k, qq, weights, attr)
 * calls (cpu/TPAc.h:19-22, cpu/PTA.h:71-76) whose results are k
 * (id, score) tuples plus the k-th score (cpu/AA.h:71-73).
 *
 * Ranking contract.  score_i = fl(...fl(fl(0 + fl(x[a0][i]*w[a0]))
 *   + fl(x[a1][i]*w[a1]))...) in the order of attr[], each product and
 *   partial sum rounded to binary32 separately (no FMA contraction), with
 *   weights indexed by attribute id (cpu/TPAc.h:129-142).  Objects rank by
 *   (score desc, id asc); the reference breaks ties by heap order, this
 *   engine by the lower object id.
 *
 * Output order.  out_ids/out_scores hold the top min(k, n) objects ranked
 *   best first (the reference's res vector is ascending, res[0] = k-th,
 *   cpu/TPAc.h:198-206; the C++ mirror in topk_amd.hpp restores that
 *   order).  Slots past min(k, n) are set to id UINT64_MAX, score -inf.
 *   *out_threshold receives the k-th (last valid) score.
 *
 * Errors.  Every entry point returns a topk_status code instead of the
 *   reference's exit(1) (input/File.h:238-241, gpu/CudaHelper.h:175-177).
 *
 * Threading.  One HIP stream per context; a context is not thread-safe.
 *   Different contexts may be used from different threads.
 *
 * Limits.  n < 2^32 (ids travel as 32 bits inside the ranking key),
 *   1 <= d <= TOPK_MAX_DIMS, 1 <= qq <= d, 1 <= k <= TOPK_MAX_K.  Input
 *   values and weights must be finite.
 */
#ifndef TOPK_H
#define TOPK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOPK_ABI_VERSION 1
#define TOPK_MAX_DIMS 32
#define TOPK_MAX_K 1024

typedef enum topk_status {
    TOPK_OK = 0,
    TOPK_ERR_INVALID = 1,      /* bad argument (k, qq, attr id, n, d, null pointer) */
    TOPK_ERR_NOMEM = 2,        /* device or pinned-host allocation failed */
    TOPK_ERR_HIP = 3,          /* a HIP runtime call failed */
    TOPK_ERR_UNSUPPORTED = 4,  /* request outside this library's limits */
    TOPK_ERR_NONFINITE = 5,    /* the store holds NaN or infinity */
    TOPK_ERR_STATE = 6,        /* e.g. topk_pta before the partitions are built */
    TOPK_ERR_NO_DEVICE = 7,    /* no usable GPU */
    TOPK_ERR_OVERFLOW = 8,     /* internal candidate buffer overflowed (retried on the exact path) */
    TOPK_ERR_FORMAT = 9,       /* a text or column file is malformed, or holds fewer rows than asked */
    TOPK_ERR_IO = 10           /* a file could not be opened, read or written */
} topk_status;

typedef struct topk_ctx topk_ctx;

/* Data distributions of the synthetic generator (the benchmark configurations). */
typedef enum topk_distr {
    TOPK_CORRELATED = 0,       /* u~U(0,1) per object, x = clip(u + N(0,0.05), 0, 1) */
    TOPK_INDEPENDENT = 1,      /* x ~ U[0,1) i.i.d. */
    TOPK_ANTICORRELATED = 2,   /* N(0,0.05) around the hyperplane sum = d/2, clipped to [0,1] */
    TOPK_MIXED = 3             /* object-range thirds: correlated, independent, anticorrelated */
} topk_distr;

/* ---- context / column store (reference: AA(n,d) + get_cdata,
 *      cpu/AA.h:66,77; File::load/gen, input/File.h:314-354) ---- */

#define TOPK_ROWMAJOR 0x1u     /* ctx_create input is row-major (n x d); transposed on device */

/* Copies an n x d FP32 table (column-major x[m*n+i] unless TOPK_ROWMAJOR)
 * into a padded column store in HBM.  The caller keeps its buffer
 * (the reference hands ownership over, cpu/AA.h:147-150). */
int topk_ctx_create(topk_ctx **out, const float *data, uint64_t n, uint32_t d, uint32_t flags);

/* Builds the column store directly in HBM with the synthetic generator
 * (same bits as topk_generate).  Replaces File::gen (input/File.h:335-354). */
int topk_ctx_create_generated(topk_ctx **out, uint64_t n, uint32_t d, int distr, uint64_t seed);

int topk_ctx_destroy(topk_ctx *ctx);

/* ---- data ingest (reference: File<T>, input/File.h) ----
 *
 * Text files hold one object per line, d values separated by ',' (the
 * "%f,%f,...,%f" rows of File::line_specifier, input/File.h:86-104).  Every
 * value converts exactly as fscanf's "%f" converts it (correctly rounded
 * binary32): the GPU parses the text, lines it cannot convert exactly on the
 * device are parsed on the host with strtof.  Blank lines are skipped (as
 * fscanf skips them); a line with another field count, a value fscanf could
 * not read, or white space before a ',' is TOPK_ERR_FORMAT (the reference's
 * loop would silently misalign the rest of the file). */

typedef struct topk_ingest_report {
    uint32_t format;        /* 0 text, 1 column file */
    uint32_t fields;        /* values per object (d) */
    uint64_t bytes;         /* file size */
    uint64_t lines;         /* text lines, blank ones included */
    uint64_t rows;          /* objects loaded (n) */
    uint64_t slow_lines;    /* lines the host parsed (exact conversion not possible on the device) */
    double ms_read;         /* file to HBM (pinned staging) */
    double ms_parse;        /* line index and conversion (text) */
    double ms_total;
} topk_ingest_report;

/* Counts the non-blank lines of a text file and the fields of its first
 * non-blank line.  Replaces File<T>::count (input/File.h:107-128), which
 * counts newlines (a last line without one, or blank lines, differ). */
int topk_csv_count(const char *path, uint64_t *rows, uint32_t *fields);

/* Reads the first n rows of d fields (0 = all rows / the first line's field
 * count) into a host n x d table, column-major out[m*n + i]
 * (File::read_scanf_t, input/File.h:264-295) or row-major with
 * TOPK_ROWMAJOR (File::read_scanf, :233-262), parsing on the current GPU.
 * rep may be NULL. */
int topk_csv_read(const char *path, uint64_t n, uint32_t d, uint32_t flags, float *out, topk_ingest_report *rep);

/* Writes an n x d table as text rows of "%f" values (File::write_printf_t,
 * input/File.h:415-430, which writes nothing for odd d; any d here). */
int topk_csv_write(const char *path, const float *data, uint64_t n, uint32_t d, uint32_t flags);

/* Binary column file (no reference counterpart): a 4096-byte header
 * ("TOPKCOL1", u32 version 1, u32 d, u64 n, u64 data offset) followed by d
 * columns of n little-endian binary32 values. */
int topk_colfile_write(const char *path, const float *data, uint64_t n, uint32_t d, uint32_t flags);
int topk_colfile_info(const char *path, uint64_t *n, uint32_t *d);

/* Builds the column store from a file: a column file streams straight into
 * HBM, a text file is uploaded and parsed on the GPU.  n / d of 0 take the
 * whole file; otherwise the first n objects (and, for a column file, the
 * first d attributes).  Replaces File(fname, ...) + File::load(get_cdata())
 * + the algorithm's init upload (validation/bench_cpu.h:156-162). flags: 0.
 * rep may be NULL. */
int topk_ctx_create_file(topk_ctx **out, const char *path, uint64_t n, uint32_t d, uint32_t flags,
                         topk_ingest_report *rep);


uint64_t topk_ctx_n(const topk_ctx *ctx);
uint32_t topk_ctx_d(const topk_ctx *ctx);

/* Copies the (possibly normalised / reordered) column store back to the
 * host as a dense column-major n x d table. */
int topk_ctx_read_columns(topk_ctx *ctx, float *out);

/* ---- init (reference: TPAc::init cpu/TPAc.h:28-32, PTA::init
 *      cpu/PTA.h:270-278, GPA::init gpu/GPA.h:62-94) ---- */

#define TOPK_INIT_NORMALIZE 0x1u   /* per-column min-max normalisation, common/common.h:29-60 */
#define TOPK_INIT_PTA       0x2u   /* build the PTA partition layout + per-partition bounds */

/* Runs the requested preprocessing; its wall time is reported as tt_init. */
int topk_init(topk_ctx *ctx, uint32_t flags);

/* Partition size used by TOPK_INIT_PTA (default 4096 objects). Must be a
 * multiple of 1024.  Call before topk_init. */
int topk_set_partition_size(topk_ctx *ctx, uint32_t objects);

/* Layout diagnostics (no reference counterpart beyond GPA::check_order,
 * gpu/GPA.h:96-114, which reads the device store back to verify it): the
 * number of partitions and of partition-order positions (incl. padding),
 * and a copy of the layout: pid[npos] (object id per position, 0xFFFFFFFF
 * for padding), poff[nparts+1], psz[nparts], pmax/pmin[nparts*d].  Null
 * outputs are skipped.  TOPK_ERR_STATE before topk_init(TOPK_INIT_PTA). */
int topk_pta_layout(topk_ctx *ctx, uint32_t *nparts, uint64_t *npos);
int topk_pta_read_layout(topk_ctx *ctx, uint32_t *pid, uint32_t *poff, uint32_t *psz, float *pmax, float *pmin);

/* reorder_attr: per object, sort its d values descending across the
 * columns, in place (gpu/reorder_attr.h:13-236; invariant
 * gpu/GPA.h:96-114).  Destroys attribute identity, as in the reference. */
int topk_reorder_attr(topk_ctx *ctx);

/* ---- queries ---- */

/* Plain full-table evaluation: TPAc::findTopK* (cpu/TPAc.h:35-358);
 * GPU analogue of the GFA stub (gpu/GFA.h:47-50). */
int topk_fte(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
             uint64_t *out_ids, float *out_scores, float *out_threshold);

/* Partition-pruned evaluation: PTA::findTopK* (cpu/PTA.h:281-671), GPA
 * (gpu/GPA.h:161-227).  Requires topk_init(ctx, TOPK_INIT_PTA).  Same
 * results as topk_fte. */
int topk_pta(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
             uint64_t *out_ids, float *out_scores, float *out_threshold);

/* Bitonic top-k aggregation: per 2048-object block top-k in LDS, then a
 * merge (gpu/BTA.h:84-255).  Same results as topk_fte. */
int topk_bta(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
             uint64_t *out_ids, float *out_scores, float *out_threshold);

/* Iterative prune + compact driven by radix select (gpu/GPA.h:161-227,
 * gpu/prune_tuples.h:90-118, gpu/radix_select.h:73-120), with bounds
 * from per-column maxima so it is exact for any weights and attribute
 * subset.  Same results as topk_fte. */
int topk_gpa(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
             uint64_t *out_ids, float *out_scores, float *out_threshold);

/* One query of a batch (cpu/TPAc.h:209-254 findTopKsimdMQ arguments). */
typedef struct topk_query {
    uint32_t k;
    uint32_t qq;
    uint8_t attr[TOPK_MAX_DIMS];      /* first qq entries used, summation order */
    float weights[TOPK_MAX_DIMS];     /* indexed by attribute id */
} topk_query;

/* Batched multi-query (validation/bench_cpu.h:198-235 IMP=3/4, over
 * TPAc::findTopKsimdMQ cpu/TPAc.h:253-358): up to 16384 queries, each with
 * its own attribute list (any order, repeats allowed), weights and k
 * (1 <= k <= k_stride <= TOPK_MAX_K).  Runs over the PTA partition layout
 * (built on first use when topk_init(TOPK_INIT_PTA) was not called): every
 * partition is loaded once per group of queries whose bound reaches the
 * query's threshold and scored for all of them.  The threshold is a ranking
 * key from a seed pass, or a score from the trained predictor once one is
 * installed (see topk_learned_train).  A second device pass in the same
 * launch re-runs the queries the first could not settle exactly: fewer than
 * k candidates (a lower threshold) or more than a candidate list holds (the
 * k-th key of the stored candidates as threshold).  A query still open after
 * it is answered by topk_fte (topk_stats.fte_fallback).  Results of
 * query q go to out_ids/out_scores[q*k_stride ...] (min(k, n) entries) and
 * out_thresholds[q]; they equal topk_fte's for the same query. */
int topk_batch(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride,
               uint64_t *out_ids, float *out_scores, float *out_thresholds);

/* Asynchronous topk_batch (no reference counterpart; for callers that keep
 * the GPU busy while they prepare the next batch or consume the last):
 * validates and stages the queries (qs may be reused when it returns),
 * launches the batch and returns its ticket.  Up to two batches are in
 * flight per context; a submit with two in flight first completes the older
 * one.  A batch's results are in its output arrays (which must stay valid
 * until then) once topk_batch_wait of its ticket or of a later one returns,
 * or once a later submit completed it. */
int topk_batch_submit(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride,
                      uint64_t *out_ids, float *out_scores, float *out_thresholds, uint64_t *ticket);

/* Completes every batch submitted up to and including ticket, oldest first;
 * topk_get_stats then describes the last batch completed.  A ticket never
 * issued is TOPK_ERR_INVALID; one already completed returns TOPK_OK. */
int topk_batch_wait(topk_ctx *ctx, uint64_t ticket);

/* Launch-shape autotuner for topk_batch (no reference counterpart; it
 * replaces the reference's compile-time block sizes, Makefile:7-54): times
 * the whole batch pipeline with HIP events on these queries for every
 * scoring shape (objects per thread, workgroup size, persistent workgroups
 * per CU), query budget per work item and select digit width, and installs
 * the fastest in the context (as topk_set_tuning would). */
typedef struct topk_batch_tuning {
    uint32_t vpt, bs, qpi, digit, grid, item;
    uint32_t variants;           /* shapes timed */
    double ms_best, ms_default;  /* device time of one batch: chosen / defaults */
    uint32_t fused;              /* 1: the query-major fused pass was chosen (batch_fused 2) */
} topk_batch_tuning;

int topk_batch_autotune(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride,
                        topk_batch_tuning *out);

/* ---- learned threshold (no reference counterpart: it replaces PTA's
 *      running threshold, cpu/PTA.h:357-422, for topk_batch) ----
 *
 * A 96 -> 64 -> 64 -> 1 ReLU MLP (FP32, 10433 parameters laid out
 * W1[64][96], b1[64], W2[64][64], b2[64], W3[64], b3) predicts a query's
 * k-th score / sum|W| from its weights, log2 k and a per-attribute quantile
 * sketch of a 1% sample of the store.  topk_batch uses (prediction - margin)
 * as the query's threshold once a predictor is installed; results stay
 * exact (a too-high threshold leaves fewer than k survivors and the query
 * is answered by topk_fte).  Stores with d <= 16 only. */
typedef struct topk_learned_config {
    uint32_t epochs;          /* passes over the training queries (default 60) */
    uint32_t batch;           /* mini-batch, a multiple of 64, at most 4096 (default 512) */
    float lr;                 /* Adam step size (default 2e-3, decayed x0.3 / x0.1 late) */
    float target_fallback;    /* fraction of calibration queries allowed to fall back (0.0005) */
    float calib_fraction;     /* held-out fraction used to calibrate the margin (0.2) */
    uint64_t seed;            /* split / init / shuffle seed (1) */
} topk_learned_config;

typedef struct topk_learned_report {
    uint32_t n_train, n_calib;
    float train_mse, calib_mse;  /* in (score / sum|W|)^2 */
    float margin;                /* installed margin, score / sum|W| units */
    float calib_fallback;        /* calibration queries whose threshold would exceed their k-th score */
    double ms_label, ms_train;   /* wall time of labelling (exact batch path) and of training */
    uint32_t saturated;          /* 1: labels reach the largest score the data allows (ties there);
                                    queries predicted near it take exact seed thresholds */
    uint32_t steps;              /* mini-batch steps run */
    double flops;                /* their arithmetic: 6 x 10433 flops per sample (forward + backward) */
    double ms_steps;             /* device time of the steps (HIP events around their launches) */
} topk_learned_report;

int topk_learned_default_config(topk_learned_config *cfg);

/* Labels the nq (>= 16) queries with the exact batch path, trains the
 * predictor on the GPU, calibrates the margin and installs it.  cfg may be
 * NULL (defaults); rep may be NULL. */
int topk_learned_train(topk_ctx *ctx, uint32_t nq, const topk_query *qs, const topk_learned_config *cfg,
                       topk_learned_report *rep);

/* Removes the installed predictor (topk_batch returns to seed thresholds). */
int topk_learned_clear(topk_ctx *ctx);

/* Reads / installs predictor parameters (10433 floats) and margin. */
int topk_learned_get(topk_ctx *ctx, float *params, float *margin);
int topk_learned_set(topk_ctx *ctx, const float *params, float margin);

/* The 96 features (x[q*96+i]) and sum|W| scales the device builds for the
 * queries; either output may be NULL. */
int topk_learned_features(topk_ctx *ctx, uint32_t nq, const topk_query *qs, float *x, float *scale);

/* Installed predictor on the queries: pred[q] = predicted k-th score,
 * tau[q] = threshold topk_batch would use.  Either output may be NULL. */
int topk_learned_predict(topk_ctx *ctx, uint32_t nq, const topk_query *qs, float *pred, float *tau);

/* The GPU forward / backward kernels on explicit rows: out[b] = MLP(x[b])
 * for nb rows of 96 features; with y given, grad = d/dparams of
 * mean_b (out[b] - y[b])^2 and *loss = that mean.  NULL outputs skipped. */
int topk_mlp_eval(topk_ctx *ctx, const float *params, const float *x, const float *y, uint32_t nb,
                  float *out, float *grad, float *loss);

/* ---- device-pointer variants (outputs stay in HBM; no host sync) ---- */

/* Enqueues topk_fte; outputs are device pointers.  Consecutive calls
 * alternate between the context's pipeline lanes (each a HIP stream with
 * its own scratch), so they may finish out of order: read outputs only after
 * topk_sync.  Errors found on the device are reported by the next
 * topk_sync. */
int topk_fte_async(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
                   uint64_t *d_ids, float *d_scores);

/* Allocation-free FTE over a caller's device-resident store.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory in the call (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No context: column m of the store is at
 * d_cols + m*ld (ld >= n, a multiple of 4; d_cols 16-byte aligned), and the
 * buffer must hold d*ld + 8192 floats, since stream tiles read up to 8191
 * values past a column's n-th (ignored).  The same results as topk_fte
 * (best first into d_ids / d_scores, k entries; d_threshold may be NULL),
 * written to device memory; *d_status (device) becomes TOPK_OK, or
 * TOPK_ERR_OVERFLOW when more than 2^20 candidates beat even the exact
 * prefix threshold (outputs then untouched).  Enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) without a host wait, a hipMalloc or
 * a memset node.  Scratch is one 11.0 MiB device array per GPU, held until
 * the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU (different streams, or threads) race.  Limits: n <
 * 2^32 - 1, k <= TOPK_MAX_K (TOPK_ERR_UNSUPPORTED beyond). */
int topk_fte_noalloc(const float *d_cols, uint64_t n, uint32_t d, uint64_t ld, uint32_t k, uint8_t qq,
                     const float *weights, const uint8_t *attr, uint64_t *d_ids, float *d_scores,
                     float *d_threshold, uint32_t *d_status, void *stream);

/* Waits for the context's stream; returns the first device-side error
 * since the last sync (TOPK_OK if none) and re-runs nothing. */
int topk_sync(topk_ctx *ctx);

/* The context's first hipStream_t (lane 0; synchronous calls run on it). */
void *topk_stream(topk_ctx *ctx);

/* ---- configuration (replaces the reference's compile-time knobs,
 *      Makefile:7-54 / debug.sh:6-91, with run-time ones) ---- */

/* Sets a launch-shape knob of the context:
 *   "stream_grid"     blocks of the FTE stream pass (0 = default)
 *   "stream_vpt"      float4 vectors per lane per column of a stream tile:
 *                     1, 2, 4 or 8 (0 = chosen from the query width)
 *   "fte_force_exact" 1 = seed FTE queries with the exact prefix pass
 *   "lanes"           1 or 2: pipelines that consecutive topk_fte_async
 *                     calls alternate between (2 overlaps one query's seed
 *                     and final passes with its neighbours' stream passes)
 *   "batch_vpt"       objects per thread of the batch scoring kernel
 *                     (d <= 8: 4, 8, 16; d <= 16: 2, 4, 8; else 2, 4)
 *   "batch_bs"        threads per batch scoring workgroup: 256 or 512
 *                     (1024 for d <= 8 with batch_vpt 4)
 *   "batch_qpi"       query budget of a batch work item (16..1024): a
 *                     partition is loaded once per item and scored for up
 *                     to this many of the queries that reach it
 *   "batch_digit"     radix digit width of the per-query select: 8 or 11
 *   "batch_grid"      persistent batch scoring workgroups
 *   "batch_item"      objects of a partition per batch work item (a multiple
 *                     of 256, 256..65536)
 *   "batch_retry"     a learned threshold that left fewer than k candidates
 *                     is rerun (second pass) with: 1 = a lower threshold
 *                     below the prediction; 2 = a seed threshold (0 = 1)
 *   "batch_retry_width" distance of that threshold below the prediction,
 *                     in calibrated margins: 1..64 (0 = default, 3)
 *   "batch_graph"     1 = capture each batch shape in a hipGraph and replay
 *   "batch_mode"      0 = learned threshold when trained, else seed;
 *                     1 = seed threshold; 2 = learned threshold
 *   "batch_mask"      layout of the batch's partition-mask kernel:
 *                     0 = query-major first pass, group-major second pass;
 *                     1 = query-major (one workgroup per query);
 *                     2 = group-major (one per partition group and query slice)
 *   "batch_fused"     1 = partition-major work items (mask, compaction, a
 *                     persistent scoring kernel); 2 = the query-major fused
 *                     pass: one workgroup per query bounds the groups,
 *                     partitions and chunks and scores the chunks that
 *                     reach the threshold, with no work list (0 = 1, or the
 *                     autotuner's choice)
 *   "mlp_train"       predictor training: 0 = two launches per mini-batch
 *                     step (default), 1 = one cooperative launch for all
 *                     steps (same parameters)
 *   "score_divide"    1 = real-data scores: the running score is divided by
 *                     qq after every term, each step rounded to binary32
 *                     (the reference's LD == 2 build, cpu/TPAc.h:143-146,
 *                     237-240, 291-294).  Applies to topk_fte, topk_fte_async
 *                     and topk_batch (whose queries then each run the FTE
 *                     path; topk_batch_autotune and topk_learned_train return
 *                     TOPK_ERR_UNSUPPORTED).  topk_pta, topk_bta, topk_gpa
 *                     keep plain sums, as the reference's PTA and GPU code
 *                     do.  0 = plain sums (default)
 * (0 for a batch knob = default or the autotuner's choice.)
 * Unknown names or values return TOPK_ERR_INVALID. */
int topk_set_tuning(topk_ctx *ctx, const char *name, int64_t value);

/* ---- statistics (reference STATS_EFF counters, cpu/AA.h:194-223) ---- */

typedef struct topk_stats {
    uint64_t tuple_count;        /* objects scored by the last query (sum over a batch) */
    uint64_t candidates;         /* keys handed to the final exact select */
    uint64_t partitions_scored;  /* PTA: partitions whose objects were scored */
    uint64_t partitions_total;
    uint32_t fallback;           /* queries of the last call whose first threshold failed: FTE
                                    queries re-run on the exact path; batch queries whose
                                    (predicted) threshold left fewer than k candidates, re-run
                                    exactly, or that topk_fte answered */
    uint32_t queries;            /* queries in the last call */
    double ms_last;              /* host wall time of the last query call */
    double tt_init;              /* host wall time of the last topk_init, ms */
    uint32_t retried;            /* batch: queries re-run with a tightened threshold */
    uint32_t items;              /* batch: work items (partition loads) of the first pass */
    uint32_t fte_fallback;       /* batch: queries neither device pass settled, answered by topk_fte */
    uint64_t terms;              /* (object, query, attribute) products scored, counted on the device
                                    (2 flops each: a multiply and an add) */
} topk_stats;

int topk_get_stats(topk_ctx *ctx, topk_stats *out);

/* Kernel timing with HIP events on the context's stream.  When enabled,
 * every launch of the named kernel family is bracketed by events; the
 * accumulated time is read by topk_timing_get (which synchronises). */
int topk_timing_enable(topk_ctx *ctx, int on);
int topk_timing_get(topk_ctx *ctx, const char *kernel, double *total_ms, uint64_t *launches);
int topk_timing_reset(topk_ctx *ctx);

/* ---- stores larger than HBM (reference: GPAm, gpu/GPAm.h:30-89) ----
 *
 * The column store lives in pinned host memory and each query streams it
 * through HBM in object-range chunks (only the queried columns): two chunk
 * stores alternate, the copy of the next chunk overlapping the exact FTE
 * pass over the current one; per-chunk top-k lists merge by ranking key
 * with global ids, so results equal topk_fte on the whole store.  A batch
 * streams every chunk once for all of its queries. */

typedef struct topk_hstore topk_hstore;

typedef struct topk_stream_report {
    uint64_t chunks;       /* chunks streamed by the last call */
    uint64_t bytes;        /* bytes copied host -> HBM by the last call */
    uint64_t queries;
    double ms_total;
} topk_stream_report;

/* Allocates a pinned host store of n x d floats and two chunk stores of
 * `chunk` objects in HBM (0: 2^26; at least 8192).  Replaces GPAm::alloc
 * (gpu/GPAm.h:30-35, cudaMallocHost + device buffer). */
int topk_hstore_create(topk_hstore **out, uint64_t n, uint32_t d, uint64_t chunk);
/* The host store to fill, column-major x[m*n + i] (AA::get_cdata, cpu/AA.h:77). */
float *topk_hstore_data(topk_hstore *hs);
uint64_t topk_hstore_n(const topk_hstore *hs);
uint32_t topk_hstore_d(const topk_hstore *hs);
int topk_hstore_destroy(topk_hstore *hs);
/* One query over the whole host store (GPAm::findTopK, gpu/GPAm.h:38-89);
 * outputs as topk_fte.  TOPK_ERR_NONFINITE if a streamed column holds NaN
 * or infinity. */
int topk_hstore_fte(topk_hstore *hs, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
                    uint64_t *out_ids, float *out_scores, float *out_threshold);
/* Many queries, each chunk streamed once for all of them; outputs as topk_batch. */
int topk_hstore_batch(topk_hstore *hs, uint32_t nq, const topk_query *qs, uint32_t k_stride, uint64_t *out_ids,
                      float *out_scores, float *out_thresholds);
int topk_hstore_report(const topk_hstore *hs, topk_stream_report *rep);

/* ---- utilities ---- */

/* Deterministic synthetic generator (replaces the time-seeded
 * input/randdataset-1.1.0 generators, randdataset.c:325-533).  Writes
 * objects [i0, i0+count) of the n-object dataset into out[m*ld + (i-i0)].
 * Counter-based: any object range reproduces the same bits. */
int topk_generate(float *out, uint64_t n, uint32_t d, uint64_t ld, int distr, uint64_t seed,
                  uint64_t i0, uint64_t count, int threads);

/* Radix select over device-resident 32-bit keys: the element of ascending
 * rank r (0-based), as gpu/radix_select.h:73-120 radix_select_gpu_findK. */
int topk_radix_select_u32(topk_ctx *ctx, const uint32_t *d_keys, uint64_t n, uint64_t r,
                          uint32_t *out);

const char *topk_strerror(int status);
int topk_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TOPK_H */
