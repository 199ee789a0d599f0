// Host-side context of the top-k engine and the kernel launch interfaces
// shared between translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/topk.h"

namespace topk {

// Column pointers and weights of one query, passed by value as a kernel
// argument (no host->device copy per query).
struct QCols {
    const float *col[TOPK_MAX_DIMS];
    float w[TOPK_MAX_DIMS];
    uint32_t m;
};

// Device-side state of one FTE / PTA query pipeline.
struct FteState {
    unsigned long long tau;   // threshold key published by the prefix pass
    unsigned int count;       // candidates appended (may exceed the capacity)
    unsigned int arrive;      // arrival ticket of the prefix pass (last block merges)
    unsigned int status;      // sticky device error bits (STATUS_*)
    unsigned int scored;      // PTA: partitions scored
    unsigned int arrive2;     // arrival ticket of a second last-block merge
    unsigned int pad;
};

enum : unsigned int { STATUS_OVERFLOW = 1u };

// Packed result block written by the final kernel, copied to pinned host
// memory in one transfer: ids[kcap] u64 | scores[kcap] f32 | thr f32 | status u32.
struct ResultLayout {
    size_t ids_off, scores_off, thr_off, status_off, bytes;
    static ResultLayout make(uint32_t kcap) {
        ResultLayout r;
        r.ids_off = 0;
        r.scores_off = r.ids_off + sizeof(uint64_t) * kcap;
        r.thr_off = r.scores_off + sizeof(float) * kcap;
        r.status_off = r.thr_off + sizeof(float);
        r.bytes = (r.status_off + sizeof(uint32_t) + 255) & ~(size_t)255;
        return r;
    }
};

struct Timing {
    double total_ms = 0;
    uint64_t launches = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

// Partition layout of the PTA path (built by topk_init with TOPK_INIT_PTA).
struct PtaLayout {
    bool built = false;
    uint32_t psize = 4096;
    uint64_t nparts = 0;
    uint64_t ld = 0;          // column stride of pcols (nparts * psize)
    float *pcols = nullptr;   // d columns in partition order
    uint32_t *pid = nullptr;  // global object id at each position
    float *pmax = nullptr;    // nparts x d per-partition per-attribute maxima
    float *pmin = nullptr;    // nparts x d minima
};

}  // namespace topk

struct topk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0;
    uint32_t d = 0;
    uint64_t ld = 0;            // padded column stride (multiple of 4096 objects)
    float *cols = nullptr;      // d * ld floats, zero padded
    // FTE pipeline scratch
    topk::FteState *st = nullptr;
    uint64_t *blk = nullptr;    // prefix-pass block candidates (256 * TOPK_MAX_K keys)
    uint64_t *cand = nullptr;   // candidate keys
    uint32_t cand_cap = 0;
    uint64_t *cand_big = nullptr;  // fallback candidate buffer (lazily sized to n)
    uint64_t cand_big_cap = 0;
    // results
    unsigned char *d_res = nullptr;
    unsigned char *h_res = nullptr;  // pinned
    topk::ResultLayout res;
    // batch scratch (grown on demand)
    void *batch_buf = nullptr;
    size_t batch_bytes = 0;
    void *h_batch = nullptr;         // pinned staging for batch results
    size_t h_batch_bytes = 0;
    // misc
    uint32_t sticky_status = 0;
    topk_stats stats{};
    bool timing_on = false;
    std::map<std::string, topk::Timing> timing;
    topk::PtaLayout pta;
    // learned predictor (owned by learned.hip)
    void *learned = nullptr;
};

namespace topk {

// Launch helpers implemented in the .hip translation units.
int fte_enqueue(topk_ctx *ctx, uint32_t k, const QCols &q, uint64_t *d_ids, float *d_scores,
                float *d_thr, uint32_t *d_status, bool big);
int build_qcols(const topk_ctx *ctx, const float *cols, uint64_t ld, uint8_t qq, const float *w,
                const uint8_t *attr, QCols &q);

// Releases the learned predictor state of a context (learned.hip).
void learned_free(topk_ctx *ctx);

// Event bracketing for topk_timing_*.
void timing_begin(topk_ctx *ctx, const char *name, hipEvent_t &ev0);
void timing_end(topk_ctx *ctx, const char *name, hipEvent_t ev0);

int hip_status(hipError_t e);

}  // namespace topk

#define TK_HIP(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return topk::hip_status(e_); \
    } while (0)
