// Stores larger than HBM: the column store stays in pinned host memory and
// queries stream it through HBM in object-range chunks.  Counterpart of the
// reference's GPAm (gpu/GPAm.h:30-89), which keeps two columns on the
// device and uploads the next one per pruning step; here each chunk carries
// every queried column, so one pass over the store answers a query (or a
// whole batch of queries: each chunk crosses PCIe once for all of them).
//
// Two chunk stores in HBM alternate: the copy of chunk c+1 runs on its own
// stream while the exact FTE path answers the queries on chunk c.  The
// top-k of the store is the top-k of the union of the chunks' top-k lists
// (an object of the global top-k is beaten by fewer than k objects of its
// own chunk), merged on the host by ranking key with global ids, so
// results equal topk_fte on the whole store bit for bit.
#include <algorithm>
#include <chrono>
#include <vector>

#include "internal.hpp"
#include "keys.hpp"

namespace topk {

constexpr uint64_t HSTORE_DEFAULT_CHUNK = 1ull << 26;
constexpr uint64_t HSTORE_MIN_CHUNK = 8192;

}  // namespace topk

struct topk_hstore {
    uint64_t n = 0;
    uint32_t d = 0;
    uint64_t chunk = 0;
    uint64_t nchunks = 0;
    float *host = nullptr;        // pinned, column-major x[m*n + i]
    topk_ctx *buf[2] = {nullptr, nullptr};  // chunk stores in HBM (capacity `chunk` objects)
    hipStream_t copy = nullptr;
    hipEvent_t ready[2] = {nullptr, nullptr};
    topk_stream_report last{};
    uint64_t *d_ids = nullptr;    // per-chunk answers of a batch's queries (device)
    float *d_scores = nullptr;
    uint64_t d_cap = 0;           // entries of d_ids / d_scores
};

using namespace topk;

namespace {

// A context with an uninitialised (zeroed) store of n objects.
int chunk_ctx(topk_ctx **out, uint64_t n, uint32_t d) {
    topk_ctx *ctx = new topk_ctx();
    ctx->n = n;
    ctx->d = d;
    hipGetDevice(&ctx->device);
    const int rc = ctx_alloc(ctx);
    if (rc) {
        topk_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return TOPK_OK;
}

// Queues the copy of chunk c's needed columns into chunk store `slot`.
int issue_copy(topk_hstore *hs, uint64_t c, int slot, uint32_t cols_mask, uint64_t &bytes) {
    topk_ctx *X = hs->buf[slot];
    const uint64_t i0 = c * hs->chunk, cnt = std::min(hs->chunk, hs->n - i0);
    for (uint32_t a = 0; a < hs->d; a++) {
        if (!(cols_mask >> a & 1u)) continue;
        TK_HIP(hipMemcpyAsync(X->cols + (uint64_t)a * X->ld, hs->host + (uint64_t)a * hs->n + i0,
                              sizeof(float) * cnt, hipMemcpyHostToDevice, hs->copy));
        bytes += sizeof(float) * cnt;
    }
    TK_HIP(hipEventRecord(hs->ready[slot], hs->copy));
    return TOPK_OK;
}

}  // namespace

extern "C" {

int topk_hstore_create(topk_hstore **out, uint64_t n, uint32_t d, uint64_t chunk) {
    if (!out || n < 1 || n >= 0xFFFFFFFFull || d < 1 || d > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    *out = nullptr;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) return TOPK_ERR_NO_DEVICE;
    if (chunk == 0) chunk = HSTORE_DEFAULT_CHUNK;
    if (chunk < HSTORE_MIN_CHUNK) return TOPK_ERR_INVALID;
    topk_hstore *hs = new topk_hstore();
    hs->n = n;
    hs->d = d;
    hs->chunk = std::min(chunk, n);
    hs->nchunks = (n + hs->chunk - 1) / hs->chunk;
    int rc = TOPK_OK;
    hipError_t e = hipHostMalloc(&hs->host, sizeof(float) * n * d, hipHostMallocDefault);
    if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc(store)", __FILE__, __LINE__);
    for (int s = 0; s < 2 && !rc; s++)
        if (s == 0 || hs->nchunks > 1) rc = chunk_ctx(&hs->buf[s], hs->chunk, d);
    if (!rc && (e = hipStreamCreateWithFlags(&hs->copy, hipStreamNonBlocking)) != hipSuccess)
        rc = hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    for (int s = 0; s < 2 && !rc; s++)
        if ((e = hipEventCreateWithFlags(&hs->ready[s], hipEventDisableTiming)) != hipSuccess)
            rc = hip_fail(e, "hipEventCreateWithFlags", __FILE__, __LINE__);
    if (rc) {
        topk_hstore_destroy(hs);
        return rc;
    }
    *out = hs;
    return TOPK_OK;
}

float *topk_hstore_data(topk_hstore *hs) { return hs ? hs->host : nullptr; }
uint64_t topk_hstore_n(const topk_hstore *hs) { return hs ? hs->n : 0; }
uint32_t topk_hstore_d(const topk_hstore *hs) { return hs ? hs->d : 0; }

int topk_hstore_destroy(topk_hstore *hs) {
    if (!hs) return TOPK_ERR_INVALID;
    if (hs->copy) (void)hipStreamSynchronize(hs->copy);
    for (int s = 0; s < 2; s++) {
        if (hs->buf[s]) topk_ctx_destroy(hs->buf[s]);
        if (hs->ready[s]) (void)hipEventDestroy(hs->ready[s]);
    }
    if (hs->copy) (void)hipStreamDestroy(hs->copy);
    if (hs->host) (void)hipHostFree(hs->host);
    (void)hipFree(hs->d_ids);
    (void)hipFree(hs->d_scores);
    delete hs;
    return TOPK_OK;
}

int topk_hstore_batch(topk_hstore *hs, uint32_t nq, const topk_query *qs, uint32_t k_stride, uint64_t *out_ids,
                      float *out_scores, float *out_thr) {
    if (!hs || !qs || nq < 1 || !out_ids || !out_scores) return TOPK_ERR_INVALID;
    uint32_t mask = 0;
    for (uint32_t q = 0; q < nq; q++) {
        const topk_query &Q = qs[q];
        if (Q.k < 1 || Q.k > TOPK_MAX_K || Q.k > k_stride || Q.qq < 1 || Q.qq > TOPK_MAX_DIMS)
            return TOPK_ERR_INVALID;
        for (uint32_t j = 0; j < Q.qq; j++) {
            if (Q.attr[j] >= hs->d) return TOPK_ERR_INVALID;
            mask |= 1u << Q.attr[j];
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t bytes = 0;
    uint32_t kmax = 1;
    for (uint32_t q = 0; q < nq; q++) kmax = std::max(kmax, qs[q].k);
    const uint64_t need = (uint64_t)nq * kmax;
    if (need > hs->d_cap) {
        (void)hipFree(hs->d_ids);
        (void)hipFree(hs->d_scores);
        hs->d_ids = nullptr;
        hs->d_scores = nullptr;
        hs->d_cap = 0;
        TK_HIP(hipMalloc(&hs->d_ids, sizeof(uint64_t) * need));
        TK_HIP(hipMalloc(&hs->d_scores, sizeof(float) * need));
        hs->d_cap = need;
    }
    // running best keys per query, best first
    std::vector<std::vector<uint64_t>> best(nq), merged(nq);
    std::vector<uint64_t> ids(need);
    std::vector<float> sc(need);
    int rc = issue_copy(hs, 0, 0, mask, bytes);
    if (!rc && hs->nchunks > 1) rc = issue_copy(hs, 1, 1, mask, bytes);
    for (uint64_t c = 0; c < hs->nchunks && !rc; c++) {
        const int slot = (int)(c & 1);
        topk_ctx *X = hs->buf[slot];
        const uint64_t i0 = c * hs->chunk;
        rc = hip_status(hipEventSynchronize(hs->ready[slot]));
        if (rc) break;
        X->n = std::min(hs->chunk, hs->n - i0);  // objects past n are never candidates
        X->store_gen++;
        rc = check_finite(X);
        // the chunk's queries pipelined over the chunk store's lanes; a
        // candidate overflow (reported by the sync) re-runs them exactly
        for (uint32_t q = 0; q < nq && !rc; q++)
            rc = topk_fte_async(X, qs[q].k, (uint8_t)qs[q].qq, qs[q].weights, qs[q].attr,
                                hs->d_ids + (uint64_t)q * kmax, hs->d_scores + (uint64_t)q * kmax);
        if (!rc) rc = topk_sync(X);
        if (rc == TOPK_ERR_OVERFLOW) {
            rc = TOPK_OK;
            for (uint32_t q = 0; q < nq && !rc; q++) {
                float thr;
                rc = topk_fte(X, qs[q].k, (uint8_t)qs[q].qq, qs[q].weights, qs[q].attr,
                              ids.data() + (uint64_t)q * kmax, sc.data() + (uint64_t)q * kmax, &thr);
            }
        } else if (!rc) {
            rc = hip_status(hipMemcpy(ids.data(), hs->d_ids, sizeof(uint64_t) * need, hipMemcpyDeviceToHost));
            if (!rc) rc = hip_status(hipMemcpy(sc.data(), hs->d_scores, sizeof(float) * need, hipMemcpyDeviceToHost));
        }
        for (uint32_t q = 0; q < nq && !rc; q++) {
            const topk_query &Q = qs[q];
            const uint64_t *qi = ids.data() + (uint64_t)q * kmax;
            const float *qsc = sc.data() + (uint64_t)q * kmax;
            const uint32_t got = (uint32_t)std::min<uint64_t>(Q.k, X->n);
            std::vector<uint64_t> &b = best[q], &m = merged[q];
            m.clear();
            size_t bi = 0;
            for (uint32_t j = 0; j < got || bi < b.size();) {
                if (m.size() == Q.k) break;
                const uint64_t kc = j < got ? rkey_from_okey(okey(qsc[j]), (uint32_t)(qi[j] + i0)) : 0ull;
                if (bi < b.size() && (j >= got || b[bi] > kc)) {
                    m.push_back(b[bi++]);
                } else {
                    m.push_back(kc);
                    j++;
                }
            }
            b.swap(m);
        }
        if (!rc && c + 2 < hs->nchunks) rc = issue_copy(hs, c + 2, slot, mask, bytes);
    }
    if (rc) {
        (void)hipStreamSynchronize(hs->copy);
        return rc;
    }
    for (uint32_t q = 0; q < nq; q++) {
        const std::vector<uint64_t> &b = best[q];
        uint64_t *oi = out_ids + (uint64_t)q * k_stride;
        float *os = out_scores + (uint64_t)q * k_stride;
        for (uint32_t j = 0; j < qs[q].k; j++) {
            if (j < b.size()) {
                oi[j] = rkey_id(b[j]);
                os[j] = rkey_score(b[j]);
            } else {
                oi[j] = ~0ull;
                os[j] = -INFINITY;
            }
        }
        if (out_thr) out_thr[q] = b.empty() ? -INFINITY : rkey_score(b.back());
    }
    hs->last.chunks = hs->nchunks;
    hs->last.bytes = bytes;
    hs->last.queries = nq;
    hs->last.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return TOPK_OK;
}

int topk_hstore_fte(topk_hstore *hs, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr,
                    uint64_t *out_ids, float *out_scores, float *out_thr) {
    if (!hs || !weights || !attr || qq < 1 || qq > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    topk_query q{};
    q.k = k;
    q.qq = qq;
    // weights are indexed by attribute id and read only at the queried ids
    // (as topk_fte reads them), so an array sized to max(attr) + 1 suffices
    for (uint32_t j = 0; j < qq; j++) {
        if (attr[j] >= hs->d) return TOPK_ERR_INVALID;
        q.attr[j] = attr[j];
        q.weights[attr[j]] = weights[attr[j]];
    }
    return topk_hstore_batch(hs, 1, &q, k, out_ids, out_scores, out_thr);
}

int topk_hstore_report(const topk_hstore *hs, topk_stream_report *rep) {
    if (!hs || !rep) return TOPK_ERR_INVALID;
    *rep = hs->last;
    return TOPK_OK;
}

}  // extern "C"
