// Host-side context of the top-k engine and the kernel launch interfaces
// shared between translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
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
    // Real-data scores (the reference's LD == 2 build, cpu/TPAc.h:143-146):
    // non-zero = divide the running score by this (qq) after every term.
    // Only the runtime-width kernels (M = 0) read it.
    float div;
};

// Device-side state of one FTE / PTA query pipeline.
struct FteState {
    unsigned long long tau;   // threshold key published by the seed pass
    unsigned int pre_count;   // exact mode: prefix keys the seed pass stored in pre[]
    unsigned int arrive;      // arrival ticket of the seed pass (last block merges)
    unsigned int status;      // sticky device error bits (STATUS_*)
    unsigned int scored;      // PTA: partitions scored
    unsigned int total;       // flat-gather mode: candidates gathered
    unsigned int cand_last;   // candidates seen by the last final select
    unsigned int spill;       // candidates that spilled past their block's slot region
    unsigned int pad;
};

// Candidate keys of one query.  Stream block b owns keys[b*slot, (b+1)*slot)
// and publishes how many keys >= tau it found in counts[b]; keys past its
// slot go to the shared spill region (spill_cap keys, counter st->spill).
// pre[0..pre_count) holds the exact-mode prefix keys.  The common case needs
// no global atomic: device-scope atomics on one counter serialise across the
// 8 XCDs and cost ~15% of the stream's bandwidth at a few thousand
// candidates; only clustered candidates reach the spill counter.
struct CandSet {
    uint64_t *keys;
    uint32_t *counts;
    uint32_t nblk;
    uint32_t slot;
    const uint64_t *pre;
    uint64_t *spill;
    uint32_t spill_cap;
};

enum : unsigned int { STATUS_OVERFLOW = 1u, STATUS_RETRY = 2u };

// FTE pipeline modes: sample-seeded (fast path), exact-prefix-seeded with
// the regular candidate buffer, and exact with an n-entry buffer.
enum { FTE_SAMPLE = 0, FTE_EXACT = 1, FTE_EXACT_BIG = 2 };

constexpr uint32_t MAX_STREAM_GRID = 4096;

// Object ids.  Ranking keys carry a 32-bit id (0xFFFFFFFF is reserved), so a
// store of 2^32 - 1 objects or more runs the FTE path per id shard of
// ID_SHARD objects (local ids) and merges the shards' top-k lists by score
// and 64-bit global id.  Stores hold at most MAX_OBJECTS objects.
constexpr uint64_t ID_LIMIT = 0xFFFFFFFFull;
constexpr uint64_t ID_SHARD = 1ull << 31;
constexpr uint64_t MAX_OBJECTS = 1ull << 38;
constexpr uint32_t MAX_ID_SHARDS = (uint32_t)(MAX_OBJECTS / ID_SHARD);
inline uint32_t id_shards(uint64_t n) { return n >= ID_LIMIT ? (uint32_t)((n + ID_SHARD - 1) / ID_SHARD) : 1u; }

// Launch-shape knobs (set by topk_set_tuning / the autotuner).
struct Tuning {
    uint32_t stream_grid = 0;   // 0: default grid cap
    int stream_vpt = 0;         // float4 vectors per lane per column in a stream tile (0: by width)
    int fte_force_exact = 0;    // 1: skip the sample seed (testing / comparison)
    int lanes = 2;              // pipeline lanes used by asynchronous queries
    // batched multi-query (0 = let the autotuner / defaults decide)
    int batch_vpt = 0;          // objects per thread of the scoring kernel
    int batch_bs = 0;           // threads per scoring workgroup (256 / 512)
    int batch_qpi = 0;          // queries per work item (16 / 32 / 64)
    int batch_digit = 0;        // radix digit width of the per-query select (8 / 11)
    int batch_grid = 0;         // persistent scoring workgroups
    int batch_item = 0;         // objects per work item
    int batch_retry = 0;        // learned misses rerun with: 0 default, 1 a lower threshold, 2 a seed
    int batch_retry_width = 0;  // the lower threshold's distance below the prediction, in margins (0: 3)
    int batch_graph = 1;        // replay captured hipGraphs
    int batch_mode = 0;         // 0 auto (learned when trained), 1 seed threshold, 2 learned threshold
    int batch_mask = 0;         // mask layout: 0 auto, 1 query-major, 2 group-major
    int batch_fused = 0;        // 0 default (work items), 1 work items, 2 query-major fused pass
    int batch_slices = 0;       // query-major first pass: workgroups per query (0: 1)
    int mlp_train = 0;          // predictor training: 0 two launches per step, 1 one cooperative launch
};

// One query of a batch as the device kernels read it (prepared on the host;
// tau is written on the device by the seed pass or the learned predictor).
// wa opens the record and (tau, tau_id, ascending, m) share one 16-byte
// word, so a kernel reads them with 16-byte loads.
struct alignas(16) BatchQuery {
    // Ascending form: when attr[] is strictly ascending, the score equals
    // the sum over all attributes a in ascending order of x[a] * wa[a] with
    // wa[a] = 0 for attributes not queried (x * 0 is +-0 and s + (+-0) == s
    // because a score is never -0).
    float wa[TOPK_MAX_DIMS];
    float w[TOPK_MAX_DIMS];         // weight of term j
    float tau;                      // objects scoring >= tau are candidates
    // (tau, tau_id) is a ranking-key threshold: an object scoring exactly tau
    // is a candidate only when its id <= tau_id (0xFFFFFFFF admits every tie),
    // so a threshold taken from a known key keeps ties past it out.
    uint32_t tau_id;
    uint32_t ascending;
    uint32_t m;                     // terms
    // Lower threshold for a rerun when tau left fewer than k candidates
    // (written with a predicted tau; -inf: rerun with a seed threshold).
    float tau_retry;
    uint32_t k;                     // min(k, n)
    uint32_t kuser;                 // k as requested
    uint32_t seed;                  // learned mode: take a seed threshold (set by the predictor)
    // The queried attributes (first 16) sorted by weight, heaviest first
    // (ties by lower id): the order a bound's fractional knapsack takes.
    float ws[16];
    uint8_t as[16];
    uint8_t attr[TOPK_MAX_DIMS];    // term j reads attribute attr[j]
    // Largest score the store allows: the binary32 score, in the query's term
    // order, of the corner of the store's box that each weight's sign prefers.
    // Rounding is monotone, so no object scores above it.  While tau ==
    // smax, an object is a candidate only if its id <= tau_id, so a
    // partition or chunk whose smallest id exceeds tau_id holds none.
    float smax;
    uint32_t pad_[3];
};
static_assert(sizeof(BatchQuery) % 16 == 0, "BatchQuery records stay 16-byte aligned");
static_assert(offsetof(BatchQuery, tau) % 16 == 0 && offsetof(BatchQuery, tau_id) == offsetof(BatchQuery, tau) + 4 &&
                  offsetof(BatchQuery, ascending) == offsetof(BatchQuery, tau) + 8 &&
                  offsetof(BatchQuery, m) == offsetof(BatchQuery, tau) + 12,
              "tau, tau_id, ascending, m form one 16-byte word");
static_assert(offsetof(BatchQuery, ws) % 16 == 0 && offsetof(BatchQuery, as) % 16 == 0, "16-byte loads");
static_assert(offsetof(BatchQuery, w) % 16 == 0 && offsetof(BatchQuery, attr) % 16 == 0 && TOPK_MAX_DIMS % 16 == 0,
              "the term weights and attributes are read with 16-byte loads");

// Expands CASE(M) for the query width m: M = 1..16 unrolled, 0 = runtime width.
#define TK_SWITCH_M(m, CASE)          \
    switch (m) {                      \
        case 1: CASE(1); break;       \
        case 2: CASE(2); break;       \
        case 3: CASE(3); break;       \
        case 4: CASE(4); break;       \
        case 5: CASE(5); break;       \
        case 6: CASE(6); break;       \
        case 7: CASE(7); break;       \
        case 8: CASE(8); break;       \
        case 9: CASE(9); break;       \
        case 10: CASE(10); break;     \
        case 11: CASE(11); break;     \
        case 12: CASE(12); break;     \
        case 13: CASE(13); break;     \
        case 14: CASE(14); break;     \
        case 15: CASE(15); break;     \
        case 16: CASE(16); break;     \
        default: CASE(0); break;      \
    }

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

// One FTE pipeline: a stream plus the per-query scratch its kernels use.
struct Lane {
    hipStream_t stream = nullptr;
    FteState *st = nullptr;
    uint64_t *blk = nullptr;     // seed-pass block winners (256 * TOPK_MAX_K keys)
    uint64_t *pre = nullptr;     // exact-mode prefix keys (TOPK_MAX_K)
    uint32_t *counts = nullptr;  // per-stream-block candidate counts (MAX_STREAM_GRID)
    uint64_t *cand = nullptr;    // candidate slots (cand_cap keys)
    uint64_t *spill = nullptr;   // shared spill region (spill_cap keys)
    uint64_t *sh_ids = nullptr;  // stores of 2^32 - 1 objects or more: per id shard, its top-k
    float *sh_scores = nullptr;  //   (local ids, TOPK_MAX_K entries per shard)
};
constexpr int NUM_LANES = 2;

struct Timing {
    double total_ms = 0;
    uint64_t launches = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

// Partition layout of the PTA path (built by topk_init with TOPK_INIT_PTA).
// Objects are grouped by a k-d tree into nparts partitions of about psize
// objects; partition p occupies positions [poff[p], poff[p] + psz[p]) of the
// partition-order columns, padded to a multiple of PART_ALIGN positions.
// Batches keep a query's partition-group list and a group's partition
// counters in LDS (4096 groups of up to 64 partitions): a layout has at most
// PTA_MAX_PARTS partitions, so stores past 2^30 objects get partitions
// larger than the requested size.
constexpr uint32_t PTA_MAX_PARTS = 1u << 18;

struct PtaLayout {
    bool built = false;
    uint32_t psize = 4096;       // target objects per partition (grown to keep PTA_MAX_PARTS)
    uint32_t nparts = 0;
    uint32_t nonempty = 0;       // partitions holding at least one object
    uint32_t maxsz = 0;          // largest partition
    uint64_t generation = 0;     // bumped whenever the layout is freed (rebuilt or released)
    uint64_t npos = 0;           // positions incl. padding (poff[nparts])
    uint64_t ld = 0;             // column stride of pcols
    float *pcols = nullptr;      // d columns in partition order, + one stream tile of padding
    uint32_t *pid = nullptr;     // global object id per position (0xFFFFFFFF: padding)
    uint32_t *cminid = nullptr;  // nchunks: smallest object id of each chunk (0xFFFFFFFF: none)
    uint32_t *poff = nullptr;    // nparts + 1 start positions
    uint32_t *psz = nullptr;     // real objects per partition
    float *pmax = nullptr;       // nparts x d per-attribute maxima
    float *pmin = nullptr;       // nparts x d per-attribute minima
    float *psum = nullptr;       // nparts: max over objects of the sum of its d attributes (rounded up)
    float *pminsum = nullptr;    // nparts: sum of the d minima (rounded down)
    uint64_t *pbound = nullptr;  // query scratch: bound key per partition
    uint64_t *bmax = nullptr;    // query scratch: max bound key per 256 partitions
    uint32_t *pcut = nullptr;    // query scratch: chunks to score per partition (0: ruled out)
    uint32_t *wchunks = nullptr; // query scratch: chunks to score per cut workgroup
    uint2 *wstat = nullptr;      // query scratch: partitions and objects to score per cut workgroup
    int cus = 0;                 // compute units of the device (the scoring grid)
    uint32_t *h_report = nullptr; // pinned, written by the query: candidates, partitions, objects (u64)
    // Bound tree: an implicit binary tree over partition ids padded to p2 (a
    // power of two); node v has children 2v, 2v+1 and leaf p is node p2 + p.
    // Leaves are numbered in k-d tree order, so a node is a contiguous,
    // spatially grouped range of partitions (exactly a k-d subtree when the
    // partition count is a power of two).  Per node: the union box, the
    // largest attribute sum, and the sum of the box minima; empty nodes have
    // tsum = -inf.  tmax / tmin are attribute-major: attribute a of node v
    // at [a * 2*p2 + v], so a wave bounding consecutive nodes reads
    // consecutive words.
    uint32_t p2 = 0;
    float *tmax = nullptr;       // d x 2*p2
    float *tmin = nullptr;       // d x 2*p2
    float *tsum = nullptr;       // 2*p2
    float *tminsum = nullptr;    // 2*p2
    uint32_t *tminid = nullptr;  // 2*p2: smallest object id under each node (0xFFFFFFFF: none)
    // Extreme-value bitmaps (built on demand, pta_extreme_bits): bit i of
    // word w of row a says whether object 64 w + i holds the store's largest
    // value of attribute a (rows 0..d-1) or its smallest (rows d..2d-1).
    uint64_t *xbits = nullptr;
    uint64_t nw = 0;             // words per row: ceil(n / 64) rounded up to 16 (128-byte rows; the padding is 0)
    // Chunk statistics: positions [c * PTA_CHUNK, (c + 1) * PTA_CHUNK) form
    // chunk c (partitions start on chunk boundaries); the crow = 2d + 2
    // statistics of chunk c -- maxima, minima, largest attribute sum (rounded
    // up; -inf: no objects) and sum of minima (rounded down) of chunk c and
    // every later chunk of its partition (suffix statistics: the bounds never
    // increase along a partition) -- are stored attribute-major, element e
    // at cstat[e * nchunks + c], so the lanes of a wave bounding consecutive
    // chunks read consecutive words.  Objects are ordered by attribute sum
    // inside a partition, so the leading chunks hold the large sums.
    uint64_t nchunks = 0;
    uint32_t crow = 0;
    float *cstat = nullptr;
    double tt_build_ms = 0;
};
constexpr uint32_t PTA_CHUNK = 256;
constexpr uint32_t PART_ALIGN = 256;
static_assert(PART_ALIGN % PTA_CHUNK == 0, "partitions start on chunk boundaries");

}  // namespace topk

struct topk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0;
    uint32_t d = 0;
    uint64_t ld = 0;            // padded column stride (multiple of 8192 objects)
    float *cols = nullptr;      // d * ld + 8192 floats, zero padded (a stream tile never needs a bounds check)
    // FTE pipeline lanes: queries alternate between them so one query's seed
    // and final passes overlap its neighbours' stream passes
    topk::Lane lane[topk::NUM_LANES > 2 ? topk::NUM_LANES : 2];
    uint32_t next_lane = 0;
    uint64_t cand_cap = 0;      // slots per lane
    uint32_t spill_cap = 0;     // spill keys per lane
    uint32_t *offsets = nullptr;   // flat-gather offsets (MAX_STREAM_GRID + 1)
    uint64_t *cand_big = nullptr;  // fallback slots, sized so no block can overflow
    uint64_t cand_big_cap = 0;
    uint64_t *flat = nullptr;   // fallback: candidates gathered contiguously
    uint64_t flat_cap = 0;
    // results
    unsigned char *d_res = nullptr;
    unsigned char *h_res = nullptr;  // pinned
    topk::ResultLayout res;
    // batched multi-query state (owned by batch.hip)
    void *batch = nullptr;
    // radix select / BTA / GPA scratch and column statistics (owned by bta_gpa.hip)
    void *aux = nullptr;
    uint64_t store_gen = 0;     // bumped whenever the column store's values change
    // misc
    topk::FteState *h_state = nullptr;  // pinned mirror of *st
    topk_stats stats{};
    bool timing_on = false;
    topk::Tuning tune;
    // FTE scores divide by qq after every term (topk_set_tuning "score_divide")
    bool score_div = false;
    std::map<std::string, topk::Timing> timing;
    topk::PtaLayout pta;
    // learned predictor (owned by learned.hip)
    void *learned = nullptr;
};

namespace topk {

// Where a final select writes its results (device pointers; thr/status may be null).
struct FinalOut {
    uint64_t *ids;
    float *scores;
    float *thr;
    uint32_t *status;
    uint32_t *cand;  // optional: the candidate count the final select saw
};

// Launch helpers implemented in the .hip translation units.
// Enqueues the final select over a candidate set on lane L (fte.hip).
void final_enqueue(topk_ctx *ctx, Lane &L, const CandSet &cs, uint32_t k, uint32_t kreq, const FinalOut &fo,
                   bool flat_fallback);
// One FTE pipeline run over objects [0, n) of the columns q points at (n = 0:
// the whole store; ids are local to those columns).
int fte_enqueue(topk_ctx *ctx, Lane &L, uint32_t k, const QCols &q, uint64_t *d_ids, float *d_scores,
                float *d_thr, uint32_t *d_status, int mode, uint64_t n = 0);
// Merges the top-k lists of `shards` id shards (shard s: sh_ids / sh_scores
// + s * TOPK_MAX_K, min(k, objects of the shard) entries, best first) into
// fo with global ids, on stream s (fte.hip).
void shard_merge_enqueue(hipStream_t s, const uint64_t *sh_ids, const float *sh_scores, uint64_t n, uint32_t shards,
                         uint32_t k, const FinalOut &fo);
int build_qcols(const topk_ctx *ctx, const float *cols, uint64_t ld, uint8_t qq, const float *w,
                const uint8_t *attr, QCols &q);

// Releases the learned predictor state of a context (learned.hip).
void learned_free(topk_ctx *ctx);
// Learned threshold (learned.hip): true when a trained predictor is
// installed; learned_tau_enqueue writes q[i].tau for i < nq on stream s
// (kernels only, so it can be captured in a graph).
bool learned_ready(const topk_ctx *ctx);
// The installed predictor marks queries for seed thresholds (saturated store).
bool learned_saturated(const topk_ctx *ctx);
uint64_t learned_version(const topk_ctx *ctx);
// Validates one query and converts it to the device form (tau left at 0).
int make_batch_query(const topk_ctx *ctx, const topk_query &q, BatchQuery &out);
// Completes the batches in flight (before the layout or the predictor changes).
void batch_quiesce(topk_ctx *ctx);
int learned_tau_enqueue(topk_ctx *ctx, hipStream_t s, BatchQuery *q, uint32_t nq);
// Releases the batch buffers and captured graphs of a context (batch.hip).
void batch_free(topk_ctx *ctx);
// Releases the radix select / BTA / GPA scratch (bta_gpa.hip).
void aux_free(topk_ctx *ctx);
int pta_build(topk_ctx *ctx);
void pta_release(topk_ctx *ctx);
int pta_extreme_bits(topk_ctx *ctx);

// Event bracketing for topk_timing_*.
void timing_begin(topk_ctx *ctx, hipStream_t s, hipEvent_t &ev0);
void timing_end(topk_ctx *ctx, hipStream_t s, const char *name, hipEvent_t ev0);

int hip_status(hipError_t e);
// hip_status, naming the failed call on stderr when TOPK_DEBUG is set.
int hip_fail(hipError_t e, const char *call, const char *file, int line);

// Allocates the column store and query scratch of a context whose n and d
// are set; rejects stores holding NaN or infinity (api.hip).
int ctx_alloc(topk_ctx *ctx);
int check_finite(topk_ctx *ctx);
// Queues the NaN / infinity check of the store; *flag (device) is or-ed with 1.
void enqueue_finite_check(const topk_ctx *ctx, hipStream_t s, unsigned int *flag);

}  // namespace topk

#define TK_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) return topk::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)
