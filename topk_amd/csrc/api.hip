// C ABI of the top-k engine: context / column store, synchronous and
// asynchronous FTE queries, statistics, kernel timing, generator.
//
// Reference interfaces replaced: AA<T,Z>(n,d) + get_cdata + ~AA
// (cpu/AA.h:63-150), File::load/gen (input/File.h:314-354),
// TPAc::findTopK* (cpu/TPAc.h:19-22), AA::benchmark counters
// (cpu/AA.h:194-223).
#include <string.h>

#include <chrono>
#include <string>
#include <cmath>
#include <thread>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace topk {

void pta_release(topk_ctx *ctx);

int hip_status(hipError_t e) {
    if (e == hipSuccess) return TOPK_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return TOPK_ERR_NOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return TOPK_ERR_NO_DEVICE;
    return TOPK_ERR_HIP;
}

int hip_fail(hipError_t e, const char *call, const char *file, int line) {
    if (getenv("TOPK_DEBUG")) fprintf(stderr, "topk %s:%d %s: %s\n", file, line, call, hipGetErrorString(e));
    return hip_status(e);
}

void timing_begin(topk_ctx *ctx, hipStream_t s, hipEvent_t &ev0) {
    ev0 = nullptr;
    if (!ctx->timing_on) return;
    if (hipEventCreate(&ev0) != hipSuccess) { ev0 = nullptr; return; }
    hipEventRecord(ev0, s);
}

void timing_end(topk_ctx *ctx, hipStream_t s, const char *name, hipEvent_t ev0) {
    if (!ev0) return;
    hipEvent_t ev1;
    if (hipEventCreate(&ev1) != hipSuccess) { hipEventDestroy(ev0); return; }
    hipEventRecord(ev1, s);
    ctx->timing[name].pending.push_back({ev0, ev1});
}

static void timing_collect(topk_ctx *ctx) {
    for (auto &kv : ctx->timing) {
        for (auto &pr : kv.second.pending) {
            float ms = 0;
            hipEventSynchronize(pr.second);
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) kv.second.total_ms += ms;
            kv.second.launches++;
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
        kv.second.pending.clear();
    }
}

int build_qcols(const topk_ctx *ctx, const float *cols, uint64_t ld, uint8_t qq, const float *w,
                const uint8_t *attr, QCols &q) {
    if (!w || !attr) return TOPK_ERR_INVALID;
    if (qq < 1 || qq > ctx->d) return TOPK_ERR_INVALID;
    memset(&q, 0, sizeof(q));
    for (uint32_t j = 0; j < qq; j++) {
        if (attr[j] >= ctx->d) return TOPK_ERR_INVALID;
        const float wj = w[attr[j]];
        if (!std::isfinite(wj)) return TOPK_ERR_INVALID;
        q.col[j] = cols + (uint64_t)attr[j] * ld;
        q.w[j] = wj;
    }
    q.m = qq;
    return TOPK_OK;
}

// ---------------------------------------------------------------------------
// store kernels
// ---------------------------------------------------------------------------

// Grid-stride over objects: a dispatch holds fewer than 2^32 work-items.
__global__ void transpose_rows_kernel(const float *__restrict__ rows, uint64_t n, uint32_t d, uint64_t ld,
                                      float *__restrict__ cols) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t m = 0; m < d; m++) cols[(uint64_t)m * ld + i] = rows[i * d + m];
}

__global__ void generate_kernel(uint64_t n, uint32_t d, uint64_t ld, int distr, uint64_t seed,
                                float *__restrict__ cols) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float v[TOPK_MAX_DIMS];
        gen_object(v, n, d, distr, seed, i);
        for (uint32_t m = 0; m < d; m++) cols[(uint64_t)m * ld + i] = v[m];
    }
}

// Sets *flag if any of the n values of column blockIdx.y is NaN or
// infinite: 16-byte loads over whole quads (columns start 32 KiB aligned),
// then the tail.
__global__ __launch_bounds__(256) void nonfinite_kernel(const float *__restrict__ cols, uint64_t n, uint64_t ld,
                                                        unsigned int *__restrict__ flag) {
    const float *c = cols + (uint64_t)blockIdx.y * ld;
    const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x, stride = (uint64_t)gridDim.x * 256;
    const uint64_t n4 = n / 4;
    bool bad = false;
    for (uint64_t i = gid; i < n4; i += stride) {
        const float4 v = reinterpret_cast<const float4 *>(c)[i];
        bad |= !isfinite(v.x) || !isfinite(v.y) || !isfinite(v.z) || !isfinite(v.w);
    }
    for (uint64_t i = n4 * 4 + gid; i < n; i += stride) bad |= !isfinite(c[i]);
    if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
}

void enqueue_finite_check(const topk_ctx *ctx, hipStream_t s, unsigned int *flag) {
    hipLaunchKernelGGL(nonfinite_kernel, dim3(512, ctx->d), dim3(256), 0, s, ctx->cols, ctx->n, ctx->ld, flag);
}

// Workgroups for one object per work-item, capped for the grid-stride kernels.
static uint32_t grid_for(uint64_t n, uint32_t bs) {
    const uint64_t g = (n + bs - 1) / bs;
    return (uint32_t)(g < (1u << 20) ? g : (1u << 20));
}

// Allocates the column store and per-query scratch of a fresh context.
int ctx_alloc(topk_ctx *ctx) {
    ctx->ld = ((ctx->n + 8191) / 8192) * 8192;
    TK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const size_t col_floats = ctx->ld * ctx->d + 8192;
    TK_HIP(hipMalloc(&ctx->cols, sizeof(float) * col_floats));
    TK_HIP(hipMemsetAsync(ctx->cols, 0, sizeof(float) * col_floats, ctx->stream));
    ctx->cand_cap = 1u << 21;
    ctx->spill_cap = 1u << 21;
    ctx->lane[0].stream = ctx->stream;
    for (int i = 0; i < NUM_LANES; i++) {
        Lane &L = ctx->lane[i];
        if (!L.stream) TK_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        TK_HIP(hipMalloc(&L.st, sizeof(FteState)));
        TK_HIP(hipMemsetAsync(L.st, 0, sizeof(FteState), L.stream));
        TK_HIP(hipMalloc(&L.blk, sizeof(uint64_t) * 256 * TOPK_MAX_K));
        TK_HIP(hipMalloc(&L.pre, sizeof(uint64_t) * TOPK_MAX_K));
        TK_HIP(hipMalloc(&L.counts, sizeof(uint32_t) * MAX_STREAM_GRID));
        TK_HIP(hipMalloc(&L.cand, sizeof(uint64_t) * ctx->cand_cap));
        TK_HIP(hipMalloc(&L.spill, sizeof(uint64_t) * ctx->spill_cap));
        const uint32_t S = id_shards(ctx->n);
        if (S > 1) {
            TK_HIP(hipMalloc(&L.sh_ids, sizeof(uint64_t) * S * TOPK_MAX_K));
            TK_HIP(hipMalloc(&L.sh_scores, sizeof(float) * S * TOPK_MAX_K));
        }
        TK_HIP(hipStreamSynchronize(L.stream));
    }
    TK_HIP(hipMalloc(&ctx->offsets, sizeof(uint32_t) * (MAX_STREAM_GRID + 1)));
    ctx->res = ResultLayout::make(TOPK_MAX_K);
    TK_HIP(hipMalloc(&ctx->d_res, ctx->res.bytes));
    TK_HIP(hipHostMalloc(&ctx->h_res, ctx->res.bytes, hipHostMallocDefault));
    TK_HIP(hipHostMalloc(&ctx->h_state, sizeof(FteState), hipHostMallocDefault));
    return TOPK_OK;
}

int check_finite(topk_ctx *ctx) {
    unsigned int *flag = reinterpret_cast<unsigned int *>(ctx->d_res + ctx->res.status_off);
    TK_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), ctx->stream));
    enqueue_finite_check(ctx, ctx->stream, flag);
    TK_HIP(hipGetLastError());
    unsigned int h = 0;
    TK_HIP(hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return h ? TOPK_ERR_NONFINITE : TOPK_OK;
}

}  // namespace topk

using namespace topk;

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int check_device() {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) return TOPK_ERR_NO_DEVICE;
    return TOPK_OK;
}

extern "C" {

int topk_abi_version(void) { return TOPK_ABI_VERSION; }

const char *topk_strerror(int s) {
    switch (s) {
        case TOPK_OK: return "ok";
        case TOPK_ERR_INVALID: return "invalid argument";
        case TOPK_ERR_NOMEM: return "out of memory";
        case TOPK_ERR_HIP: return "HIP runtime error";
        case TOPK_ERR_UNSUPPORTED: return "unsupported request";
        case TOPK_ERR_NONFINITE: return "store holds NaN or infinity";
        case TOPK_ERR_STATE: return "context not initialised for this call";
        case TOPK_ERR_NO_DEVICE: return "no usable GPU";
        case TOPK_ERR_OVERFLOW: return "candidate buffer overflow";
        case TOPK_ERR_FORMAT: return "malformed file";
        case TOPK_ERR_IO: return "file input/output error";
        default: return "unknown status";
    }
}

int topk_ctx_destroy(topk_ctx *ctx) {
    if (!ctx) return TOPK_ERR_INVALID;
    for (int i = 0; i < NUM_LANES; i++)
        if (ctx->lane[i].stream) hipStreamSynchronize(ctx->lane[i].stream);
    timing_collect(ctx);
    learned_free(ctx);
    hipFree(ctx->cols);
    for (int i = 0; i < NUM_LANES; i++) {
        Lane &L = ctx->lane[i];
        hipFree(L.st);
        hipFree(L.blk);
        hipFree(L.pre);
        hipFree(L.counts);
        hipFree(L.cand);
        hipFree(L.spill);
        hipFree(L.sh_ids);
        hipFree(L.sh_scores);
        if (i > 0 && L.stream) hipStreamDestroy(L.stream);
    }
    hipFree(ctx->offsets);
    hipFree(ctx->cand_big);
    hipFree(ctx->flat);
    hipFree(ctx->d_res);
    if (ctx->h_res) hipHostFree(ctx->h_res);
    if (ctx->h_state) hipHostFree(ctx->h_state);
    batch_free(ctx);
    aux_free(ctx);
    pta_release(ctx);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return TOPK_OK;
}

int topk_ctx_create(topk_ctx **out, const float *data, uint64_t n, uint32_t d, uint32_t flags) {
    if (!out || !data || n < 1 || d < 1 || d > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    if (n > MAX_OBJECTS) return TOPK_ERR_UNSUPPORTED;
    *out = nullptr;
    int rc = check_device();
    if (rc) return rc;
    topk_ctx *ctx = new topk_ctx();
    ctx->n = n;
    ctx->d = d;
    hipGetDevice(&ctx->device);
    rc = ctx_alloc(ctx);
    if (!rc) {
        if (flags & TOPK_ROWMAJOR) {
            float *tmp = nullptr;
            if (hipMalloc(&tmp, sizeof(float) * n * d) != hipSuccess) rc = TOPK_ERR_NOMEM;
            if (!rc && hipMemcpyAsync(tmp, data, sizeof(float) * n * d, hipMemcpyHostToDevice, ctx->stream))
                rc = TOPK_ERR_HIP;
            if (!rc) {
                hipLaunchKernelGGL(transpose_rows_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, tmp, n,
                                   d, ctx->ld, ctx->cols);
                if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = TOPK_ERR_HIP;
            }
            hipFree(tmp);
        } else {
            for (uint32_t m = 0; m < d && !rc; m++)
                if (hipMemcpyAsync(ctx->cols + (uint64_t)m * ctx->ld, data + (uint64_t)m * n, sizeof(float) * n,
                                   hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                    rc = TOPK_ERR_HIP;
        }
    }
    if (!rc) rc = check_finite(ctx);
    if (rc) { topk_ctx_destroy(ctx); return rc; }
    *out = ctx;
    return TOPK_OK;
}

int topk_ctx_create_generated(topk_ctx **out, uint64_t n, uint32_t d, int distr, uint64_t seed) {
    if (!out || n < 1 || d < 1 || d > TOPK_MAX_DIMS || distr < 0 || distr > 3) return TOPK_ERR_INVALID;
    if (n > MAX_OBJECTS) return TOPK_ERR_UNSUPPORTED;
    *out = nullptr;
    int rc = check_device();
    if (rc) return rc;
    topk_ctx *ctx = new topk_ctx();
    ctx->n = n;
    ctx->d = d;
    hipGetDevice(&ctx->device);
    rc = ctx_alloc(ctx);
    if (!rc) {
        hipLaunchKernelGGL(generate_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, d, ctx->ld, distr,
                           seed, ctx->cols);
        if (hipGetLastError() != hipSuccess) rc = TOPK_ERR_HIP;
    }
    if (!rc) rc = check_finite(ctx);
    if (rc) { topk_ctx_destroy(ctx); return rc; }
    *out = ctx;
    return TOPK_OK;
}

uint64_t topk_ctx_n(const topk_ctx *ctx) { return ctx ? ctx->n : 0; }
uint32_t topk_ctx_d(const topk_ctx *ctx) { return ctx ? ctx->d : 0; }
void *topk_stream(topk_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int topk_ctx_read_columns(topk_ctx *ctx, float *out) {
    if (!ctx || !out) return TOPK_ERR_INVALID;
    for (uint32_t m = 0; m < ctx->d; m++)
        TK_HIP(hipMemcpyAsync(out + (uint64_t)m * ctx->n, ctx->cols + (uint64_t)m * ctx->ld,
                              sizeof(float) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TOPK_OK;
}

static int check_query_args(topk_ctx *ctx, uint32_t k) {
    if (!ctx) return TOPK_ERR_INVALID;
    if (k < 1) return TOPK_ERR_INVALID;
    if (k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    return TOPK_OK;
}

// Copies a finished result block to the caller (best first, sentinels past n).
static void emit_result(topk_ctx *ctx, uint32_t k, uint64_t *out_ids, float *out_scores, float *out_thr) {
    const uint64_t *ids = reinterpret_cast<const uint64_t *>(ctx->h_res + ctx->res.ids_off);
    const float *sc = reinterpret_cast<const float *>(ctx->h_res + ctx->res.scores_off);
    if (out_ids) memcpy(out_ids, ids, sizeof(uint64_t) * k);
    if (out_scores) memcpy(out_scores, sc, sizeof(float) * k);
    if (out_thr) *out_thr = *reinterpret_cast<const float *>(ctx->h_res + ctx->res.thr_off);
}

// Fallback buffers: slots large enough that no stream block can overflow
// (every block's objects fit in its slot) and a flat gather buffer.
static int ensure_big(topk_ctx *ctx, uint64_t n) {
    const uint64_t need = n + (uint64_t)MAX_STREAM_GRID * 8192 + TOPK_MAX_K;
    if (ctx->cand_big && ctx->cand_big_cap >= need) return TOPK_OK;
    hipFree(ctx->cand_big);
    hipFree(ctx->flat);
    ctx->cand_big = nullptr;
    ctx->flat = nullptr;
    TK_HIP(hipMalloc(&ctx->cand_big, sizeof(uint64_t) * need));
    ctx->cand_big_cap = need;
    TK_HIP(hipMalloc(&ctx->flat, sizeof(uint64_t) * (n + TOPK_MAX_K)));
    ctx->flat_cap = n + TOPK_MAX_K;
    return TOPK_OK;
}

// One FTE query over objects [0, n) of the columns q points at, results to
// device pointers (status in the context's result block).  The
// sample-seeded pipeline runs first; if its guess left too few candidates
// (STATUS_RETRY) or too many for the buffer (STATUS_OVERFLOW), the
// exact-prefix pipeline runs, and after an overflow there an n-entry
// candidate buffer.
static int fte_run_range(topk_ctx *ctx, uint32_t k, const QCols &q, uint64_t n, uint64_t *d_ids, float *d_sc,
                         float *d_thr) {
    uint32_t *d_status = reinterpret_cast<uint32_t *>(ctx->d_res + ctx->res.status_off);
    Lane &L = ctx->lane[0];
    int mode = ctx->tune.fte_force_exact ? FTE_EXACT : FTE_SAMPLE;
    for (int attempt = 0; attempt < 3; attempt++) {
        if (mode == FTE_EXACT_BIG) {
            int rc = ensure_big(ctx, n);
            if (rc) return rc;
        }
        int rc = fte_enqueue(ctx, L, k, q, d_ids, d_sc, d_thr, d_status, mode, n);
        if (rc) return rc;
        TK_HIP(hipMemcpyAsync(ctx->h_res, ctx->d_res, ctx->res.bytes, hipMemcpyDeviceToHost, L.stream));
        TK_HIP(hipMemcpyAsync(ctx->h_state, L.st, sizeof(FteState), hipMemcpyDeviceToHost, L.stream));
        TK_HIP(hipStreamSynchronize(L.stream));
        const uint32_t status = *reinterpret_cast<const uint32_t *>(ctx->h_res + ctx->res.status_off);
        if (!status) {
            ctx->stats.candidates += ctx->h_state->cand_last;
            return TOPK_OK;
        }
        TK_HIP(hipMemsetAsync(&L.st->status, 0, sizeof(uint32_t), L.stream));
        ctx->stats.fallback = 1;
        // a sample guess that overflowed is re-run with the exact prefix
        // seed first; the n-entry buffer only serves an exact run's overflow
        mode = (status & STATUS_OVERFLOW) && mode != FTE_SAMPLE ? FTE_EXACT_BIG : FTE_EXACT;
    }
    return TOPK_ERR_OVERFLOW;
}

// The columns of q shifted to the first object of id shard s.
static QCols shard_cols(const QCols &q, uint32_t s) {
    QCols r = q;
    for (uint32_t j = 0; j < q.m; j++) r.col[j] = q.col[j] + (uint64_t)s * ID_SHARD;
    return r;
}

// Runs one FTE query into the context's result block and copies it to the
// pinned host mirror: one run over the store, or one per id shard and a
// device merge of their top-k lists.
static int fte_run_sync(topk_ctx *ctx, uint32_t k, const QCols &q) {
    uint64_t *d_ids = reinterpret_cast<uint64_t *>(ctx->d_res + ctx->res.ids_off);
    float *d_sc = reinterpret_cast<float *>(ctx->d_res + ctx->res.scores_off);
    float *d_thr = reinterpret_cast<float *>(ctx->d_res + ctx->res.thr_off);
    uint32_t *d_status = reinterpret_cast<uint32_t *>(ctx->d_res + ctx->res.status_off);
    ctx->stats.fallback = 0;
    ctx->stats.retried = 0;
    ctx->stats.fte_fallback = 0;
    ctx->stats.candidates = 0;
    const uint32_t S = id_shards(ctx->n);
    if (S == 1) return fte_run_range(ctx, k, q, ctx->n, d_ids, d_sc, d_thr);
    Lane &L = ctx->lane[0];
    for (uint32_t s = 0; s < S; s++) {
        const uint64_t n0 = (uint64_t)s * ID_SHARD, ns = ctx->n - n0 < ID_SHARD ? ctx->n - n0 : ID_SHARD;
        const int rc = fte_run_range(ctx, k, shard_cols(q, s), ns, L.sh_ids + (uint64_t)s * TOPK_MAX_K,
                                     L.sh_scores + (uint64_t)s * TOPK_MAX_K, nullptr);
        if (rc) return rc;
    }
    shard_merge_enqueue(L.stream, L.sh_ids, L.sh_scores, ctx->n, S, k, FinalOut{d_ids, d_sc, d_thr, d_status});
    TK_HIP(hipGetLastError());
    TK_HIP(hipMemcpyAsync(ctx->h_res, ctx->d_res, ctx->res.bytes, hipMemcpyDeviceToHost, L.stream));
    TK_HIP(hipStreamSynchronize(L.stream));
    return TOPK_OK;
}

int topk_fte(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, uint64_t *out_ids,
             float *out_scores, float *out_threshold) {
    int rc = check_query_args(ctx, k);
    if (rc) return rc;
    QCols q;
    rc = build_qcols(ctx, ctx->cols, ctx->ld, qq, weights, attr, q);
    if (rc) return rc;
    if (ctx->score_div) q.div = (float)qq;
    const double t0 = now_ms();
    rc = fte_run_sync(ctx, k, q);
    if (rc) return rc;
    emit_result(ctx, k, out_ids, out_scores, out_threshold);
    ctx->stats.ms_last = now_ms() - t0;
    ctx->stats.tuple_count = ctx->n;
    ctx->stats.terms = ctx->n * qq;
    ctx->stats.queries = 1;
    return TOPK_OK;
}

int topk_fte_async(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
                   uint64_t *d_ids, float *d_scores) {
    int rc = check_query_args(ctx, k);
    if (rc) return rc;
    if (!d_ids || !d_scores) return TOPK_ERR_INVALID;
    QCols q;
    rc = build_qcols(ctx, ctx->cols, ctx->ld, qq, weights, attr, q);
    if (rc) return rc;
    if (ctx->score_div) q.div = (float)qq;
    // alternate lanes so consecutive queries overlap (timing runs serialise on lane 0)
    const uint32_t lanes = ctx->timing_on ? 1u : (uint32_t)ctx->tune.lanes;
    Lane &L = ctx->lane[ctx->next_lane % lanes];
    ctx->next_lane = (ctx->next_lane + 1) % lanes;
    const int mode = ctx->tune.fte_force_exact ? FTE_EXACT : FTE_SAMPLE;
    const uint32_t S = id_shards(ctx->n);
    if (S == 1) return fte_enqueue(ctx, L, k, q, d_ids, d_scores, nullptr, nullptr, mode);
    // per id shard into the lane's shard lists, then the merge
    for (uint32_t s = 0; s < S; s++) {
        const uint64_t n0 = (uint64_t)s * ID_SHARD, ns = ctx->n - n0 < ID_SHARD ? ctx->n - n0 : ID_SHARD;
        rc = fte_enqueue(ctx, L, k, shard_cols(q, s), L.sh_ids + (uint64_t)s * TOPK_MAX_K,
                         L.sh_scores + (uint64_t)s * TOPK_MAX_K, nullptr, nullptr, mode, ns);
        if (rc) return rc;
    }
    shard_merge_enqueue(L.stream, L.sh_ids, L.sh_scores, ctx->n, S, k, FinalOut{d_ids, d_scores, nullptr, nullptr});
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

int topk_sync(topk_ctx *ctx) {
    if (!ctx) return TOPK_ERR_INVALID;
    uint32_t any = 0;
    for (int i = 0; i < NUM_LANES; i++) {
        Lane &L = ctx->lane[i];
        TK_HIP(hipStreamSynchronize(L.stream));
        uint32_t status = 0;
        TK_HIP(hipMemcpy(&status, &L.st->status, sizeof(status), hipMemcpyDeviceToHost));
        if (status) {
            TK_HIP(hipMemset(&L.st->status, 0, sizeof(uint32_t)));
            any |= status;
        }
    }
    return any ? TOPK_ERR_OVERFLOW : TOPK_OK;
}

int topk_set_tuning(topk_ctx *ctx, const char *name, int64_t value) {
    if (!ctx || !name) return TOPK_ERR_INVALID;
    const std::string key(name);
    if (key == "stream_grid") {
        if (value < 0 || value > (int64_t)MAX_STREAM_GRID) return TOPK_ERR_INVALID;
        ctx->tune.stream_grid = (uint32_t)value;
    } else if (key == "stream_vpt") {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) return TOPK_ERR_INVALID;
        ctx->tune.stream_vpt = (int)value;
    } else if (key == "lanes") {
        if (value < 1 || value > NUM_LANES) return TOPK_ERR_INVALID;
        ctx->tune.lanes = (int)value;
    } else if (key == "fte_force_exact") {
        ctx->tune.fte_force_exact = value ? 1 : 0;
    } else if (key == "batch_vpt") {
        if (value != 0 && value != 2 && value != 4 && value != 8 && value != 16) return TOPK_ERR_INVALID;
        ctx->tune.batch_vpt = (int)value;
    } else if (key == "batch_bs") {
        if (value != 0 && value != 256 && value != 512 && value != 1024) return TOPK_ERR_INVALID;
        ctx->tune.batch_bs = (int)value;
    } else if (key == "batch_qpi") {
        if (value != 0 && (value < 16 || value > 1024)) return TOPK_ERR_INVALID;
        ctx->tune.batch_qpi = (int)value;
    } else if (key == "batch_digit") {
        if (value != 0 && value != 8 && value != 11) return TOPK_ERR_INVALID;
        ctx->tune.batch_digit = (int)value;
    } else if (key == "batch_grid") {
        if (value < 0 || value > 65536) return TOPK_ERR_INVALID;
        ctx->tune.batch_grid = (int)value;
    } else if (key == "batch_item") {
        if (value != 0 && (value < 256 || value > 65536 || value % 256)) return TOPK_ERR_INVALID;
        ctx->tune.batch_item = (int)value;
    } else if (key == "batch_retry") {
        if (value < 0 || value > 2) return TOPK_ERR_INVALID;
        ctx->tune.batch_retry = (int)value;
    } else if (key == "batch_retry_width") {
        if (value < 0 || value > 64) return TOPK_ERR_INVALID;
        ctx->tune.batch_retry_width = (int)value;
    } else if (key == "batch_graph") {
        ctx->tune.batch_graph = value ? 1 : 0;
    } else if (key == "batch_mode") {
        if (value < 0 || value > 2) return TOPK_ERR_INVALID;
        ctx->tune.batch_mode = (int)value;
    } else if (key == "batch_mask") {
        if (value < 0 || value > 2) return TOPK_ERR_INVALID;
        ctx->tune.batch_mask = (int)value;
    } else if (key == "batch_slices") {
        if (value < 0 || value > 64) return TOPK_ERR_INVALID;
        batch_quiesce(ctx);
        ctx->tune.batch_slices = (int)value;
    } else if (key == "mlp_train") {
        if (value < 0 || value > 1) return TOPK_ERR_INVALID;
        ctx->tune.mlp_train = (int)value;
    } else if (key == "batch_fused") {
        if (value < 0 || value > 2) return TOPK_ERR_INVALID;
        ctx->tune.batch_fused = (int)value;
    } else if (key == "score_divide") {
        if (value < 0 || value > 1) return TOPK_ERR_INVALID;
        batch_quiesce(ctx);  // batches in flight complete under the mode they ran with
        ctx->score_div = value != 0;
    } else {
        return TOPK_ERR_INVALID;
    }
    return TOPK_OK;
}

int topk_get_stats(topk_ctx *ctx, topk_stats *out) {
    if (!ctx || !out) return TOPK_ERR_INVALID;
    *out = ctx->stats;
    return TOPK_OK;
}

int topk_timing_enable(topk_ctx *ctx, int on) {
    if (!ctx) return TOPK_ERR_INVALID;
    ctx->timing_on = on != 0;
    return TOPK_OK;
}

int topk_timing_get(topk_ctx *ctx, const char *kernel, double *total_ms, uint64_t *launches) {
    if (!ctx || !kernel) return TOPK_ERR_INVALID;
    timing_collect(ctx);
    auto it = ctx->timing.find(kernel);
    if (total_ms) *total_ms = it == ctx->timing.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == ctx->timing.end() ? 0 : it->second.launches;
    return TOPK_OK;
}

int topk_timing_reset(topk_ctx *ctx) {
    if (!ctx) return TOPK_ERR_INVALID;
    timing_collect(ctx);
    ctx->timing.clear();
    return TOPK_OK;
}

int topk_generate(float *out, uint64_t n, uint32_t d, uint64_t ld, int distr, uint64_t seed, uint64_t i0,
                  uint64_t count, int threads) {
    if (!out || d < 1 || d > TOPK_MAX_DIMS || distr < 0 || distr > 3 || i0 + count > n || ld < count)
        return TOPK_ERR_INVALID;
    if (threads < 1) threads = (int)std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > count / 4096 + 1) threads = (int)(count / 4096 + 1);
    auto work = [&](int t) {
        const uint64_t lo = i0 + count * t / threads, hi = i0 + count * (t + 1) / threads;
        float v[TOPK_MAX_DIMS];
        for (uint64_t i = lo; i < hi; i++) {
            gen_object(v, n, d, distr, seed, i);
            for (uint32_t m = 0; m < d; m++) out[(uint64_t)m * ld + (i - i0)] = v[m];
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return TOPK_OK;
}

}  // extern "C"
