// Batched multi-query top-k on gfx950 (topk_batch).
//
// Reference: validation/bench_cpu.h:198-235 (IMP=3/4 multiple queries with
// per-query attribute subsets and k) over TPAc::findTopKsimdMQ
// (cpu/TPAc.h:253-358).  Each query's answer equals topk_fte's.
//
// The batch runs over the partition layout of the PTA path (built on first
// use).  One call is a fixed kernel sequence on one stream, captured once
// per batch shape in a hipGraph and replayed:
//
//  0. prep (one wave per query): the device form of every query (terms,
//     ascending form, LP bound arguments) and cleared per-batch counters.
//  1. threshold per query, a ranking key (tau_q, tau_id_q): objects whose
//     (score, id) key reaches it are candidates.
//       - seed mode: a beam search down the bound tree picks the query's 4
//         best-bounded partitions; the exact k-th key among their first
//         4096 objects (the leading 1024 of each: largest attribute sums)
//         is a lower bound of the query's k-th key.
//       - learned mode: an MLP predicts the k-th score from (query, data
//         sketch), lowered by a calibrated margin (learned.hip).
//  2. mask (one workgroup per query): groups of 2^H partitions whose bound
//     reaches tau_q, then their partitions; q is appended to the query list
//     of every partition whose bound reaches tau_q.
//  3. compaction (one thread per partition): each query list becomes work
//     items of at most qpi queries over object chunks; items with many
//     (object, query) pairs go to the front of the list.
//  4. scoring (persistent workgroups taking items from a global ticket):
//     a workgroup loads the chunk's objects into registers once and scores
//     them for every query of the item; survivors are appended to the
//     query's candidate list (wave-aggregated atomic).  Scores use the
//     query's own term order: a fixed ascending register walk when the terms
//     are ascending, else a wave-uniform switch on the attribute id.
//  5. final select per query (one workgroup each): radix select over the
//     candidates (in registers up to 8192, else in memory), rank sort,
//     best-first output.
//  6. a second pass of 0-5 in the same graph for the queries the first left
//     open: a predicted threshold with fewer than k candidates reruns with a
//     lower retry threshold (or a seed), and a candidate list over capacity
//     reruns with the k-th key of its stored candidates as threshold (ties
//     past that key stay out).
//
// Exactness: every object whose key reaches the threshold lies in a
// partition whose bound reaches tau_q, so it is a candidate.  With c >= k
// candidates the top k are among them.  A query still open after the second
// pass (too few candidates or too many) is answered by the plain exact FTE
// path on the host side of the call and counted in topk_stats.fallback.
#include <string.h>

#include <array>
#include <chrono>
#include <map>
#include <vector>

#include "bound.hpp"
#include "device.hpp"
#include "emit.hpp"
#include "internal.hpp"

namespace topk {

constexpr uint32_t BATCH_MAX_Q = 16384;   // queries per call
constexpr uint32_t BATCH_CAP = 32768;     // candidate keys per query
constexpr uint32_t BATCH_MAX_GRID = 65536;  // persistent scoring workgroups
// lists up to 8192 keys are selected in registers, longer ones in memory;
// 256 threads x 32 keys (127 registers) keep a 1024-query batch's
// workgroups resident at once -- 512 x 16 needed eight waves per SIMD where
// five fit, and ran in two rounds
constexpr int BFIN_BS = 256, BFIN_KPT = 32;
// 256 threads x 16 objects: a batch's 1024 seed workgroups are resident at
// once at every d (512 x 8 needed eight waves per SIMD, five fit at d = 16)
constexpr int BSEED_BS = 256, BSEED_KPT = 16;
constexpr uint32_t BSEED_T = 4;                          // seed partitions per query
// objects a seed scores (4096): the leading objects of each seed partition,
// which hold its largest attribute sums
constexpr uint32_t BSEED_OBJS = BSEED_BS * BSEED_KPT;
constexpr int BAUX_BS = 256;                             // bound / mask / compaction kernels
constexpr uint32_t ITEM_MIN_OBJECTS = 2048;              // smallest object chunk of a work item
constexpr int BMASK_BS = 256;                            // mask workgroup
constexpr uint32_t BMASK_MAX_PER = 64;                   // partitions per mask group (2^H)
constexpr uint32_t BMASK_PAIRS = 4096;                   // (query, partition) pairs a mask workgroup holds
static_assert(BMASK_PAIRS * BMASK_MAX_PER == PTA_MAX_PARTS, "every layout fits the group lists");
enum { BMODE_SEED = 1, BMODE_LEARNED = 2, BMODE_GIVEN = 3 };
// Per-query flags the first pass leaves for the second (in-graph) pass.
enum : uint32_t { QF_DONE = 0, QF_SEED = 1, QF_GIVEN = 2, QF_LOWER = 3 };
// Device counters of one batch: [0] heavy work items (front of the list),
// [1] light work items (back of the list), [2..3] scored (partition, query)
// pairs, [4..5] scored (object, query) pairs, [6] work ticket, [7] unused,
// [8..9] scored (object, query, attribute) terms.
constexpr uint32_t BCTR_WORDS = 10;

// Work item: objects [chunk * ochunk, (chunk + 1) * ochunk) of partition p,
// scored for the qcnt queries at plist[p * nq + qbeg ...].
struct BatchItem {
    uint32_t p;
    uint32_t qbeg, qcnt;
    uint32_t chunk;
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Device form of query q, one wave per query (lane j: term j / attribute
// j): terms in query order, the ascending form, and the bound arguments --
// the distinct attributes with aggregated FP64 weights, heaviest first (ties
// by attribute id).  taus (optional) gives the thresholds of a second pass.
// rmax / rmin: the store's per-attribute maxima / minima (attribute a at
// [a * rstride]).  Every thread also clears its share of the per-batch
// counters.
__global__ __launch_bounds__(64) void batch_prep_kernel(const topk_query *__restrict__ raw,
                                                        const float *__restrict__ taus, uint32_t nq, uint64_t n,
                                                        uint32_t d, const float *__restrict__ rmax,
                                                        const float *__restrict__ rmin, uint64_t rstride,
                                                        BatchQuery *__restrict__ out,
                                                        BoundArgs *__restrict__ outb, unsigned int *__restrict__ pcnt,
                                                        uint32_t L, unsigned int *__restrict__ ccount,
                                                        unsigned int *__restrict__ ctr, unsigned int *__restrict__ flist) {
    const uint32_t lane = lane_id();
    const uint64_t stride = (uint64_t)gridDim.x * 64;
    for (uint64_t i = (uint64_t)blockIdx.x * 64 + lane; i < L; i += stride) pcnt[i] = 0u;
    for (uint64_t i = (uint64_t)blockIdx.x * 64 + lane; i < nq; i += stride) ccount[i] = 0u;
    if (blockIdx.x == 0 && lane < BCTR_WORDS) ctr[lane] = 0u;
    if (blockIdx.x == 0 && lane == 0 && flist) flist[0] = 0u;
    const uint32_t q = blockIdx.x;
    if (q >= nq) return;
    const topk_query &Q = raw[q];
    BatchQuery &b = out[q];
    BoundArgs &g = outb[q];
    const uint32_t qq = Q.qq, j = lane;
    const bool term = j < qq;
    const uint32_t aj = term ? Q.attr[j] : 0u;
    const float wj = term ? Q.weights[aj] : 0.0f;
    const uint32_t aprev = __shfl(aj, j ? j - 1 : 0, 64);
    const bool asc = __ballot(term && j > 0 && aj <= aprev) == 0ull;
    if (j < TOPK_MAX_DIMS) {
        b.attr[j] = (uint8_t)aj;
        b.w[j] = wj;
    }
    // lane a: aggregated weight of attribute a
    double agg = 0.0;
    bool used = false;
    for (uint32_t t = 0; t < qq; t++) {
        const uint32_t at = __shfl(aj, t, 64);
        const float wt = __shfl(wj, t, 64);
        if (at == lane) {
            agg += (double)wt;
            used = true;
        }
    }
    if (lane < TOPK_MAX_DIMS) b.wa[lane] = asc && used ? Q.weights[lane] : 0.0f;
    uint32_t r = 0;  // rank: heavier first, ties by lower attribute id
    for (uint32_t c = 0; c < d; c++) {
        const double ac = __shfl(agg, c, 64);
        const bool uc = __shfl((int)used, c, 64) != 0;
        r += uc && (ac > agg || (ac == agg && c < lane)) ? 1u : 0u;
    }
    // the first 16 ranks, staged in LDS so that every ws / as slot is written
    // once: rank r's weight (exact for ascending queries, each attribute
    // once) and attribute, zeros past the distinct attributes
    __shared__ float sws[16];
    __shared__ uint32_t sas[16];
    if (lane < 16) {
        sws[lane] = 0.0f;
        sas[lane] = 0u;
    }
    __syncthreads();
    if (used && lane < d) {
        g.attr[r] = (uint8_t)lane;
        g.w[r] = agg;
        if (r < 16) {
            sws[r] = (float)agg;
            sas[r] = lane;
        }
    }
    __syncthreads();
    if (lane < 16) {
        b.ws[lane] = sws[lane];
        b.as[lane] = (uint8_t)sas[lane];
    }
    const uint32_t nd = (uint32_t)__popcll(__ballot(used && lane < d));
    // largest score: each term at the end of its attribute's range that its
    // weight's sign prefers (lane j: term j's product, one load), summed in
    // term order as the scores are
    const float prj = term ? (wj >= 0.0f ? rmax[aj * rstride] : rmin[aj * rstride]) * wj : 0.0f;
    float sm = 0.0f;
    for (uint32_t t = 0; t < qq; t++) sm = sm + __shfl(prj, t, 64);
    if (lane == 0) {
        b.smax = sm;
        b.tau = taus ? taus[q] : 0.0f;
        b.tau_retry = -INFINITY;
        b.tau_id = 0xFFFFFFFFu;
        b.seed = 0u;
        b.k = (uint32_t)(n < Q.k ? n : Q.k);
        b.kuser = Q.k;
        b.m = qq;
        b.ascending = asc ? 1u : 0u;
        g.d = d;
        g.nd = nd;
    }
}

// Bound tree of the partition layout as the batch kernels read it.
struct TreeArgs {
    const float *tmax, *tmin, *tsum, *tminsum;  // tmax / tmin attribute-major: [a * 2*p2 + v]
    const uint32_t *tminid;                     // smallest object id under node v
    uint32_t p2, d;
};

// Query Q's threshold sits at the largest score the store allows, so an
// object is a candidate only if its id is at most tau_id: a node, partition
// or chunk whose smallest id exceeds tau_id holds no candidate.
__device__ __forceinline__ bool at_smax(const BatchQuery &Q) {
    return Q.tau == Q.smax && Q.tau_id != 0xFFFFFFFFu;
}

// Upper bound of query b's score over tree node v (-inf for an empty node);
// NB >= b.nd.  The node's loads are issued together, not behind the
// emptiness test.
template <int NB>
__device__ __forceinline__ float node_bound(const BoundArgs &b, const TreeArgs &t, uint32_t v) {
    const float ts = t.tsum[v];
    const float ub = lp_bound<NB>(b, t.tmax + v, t.tmin + v, ts, t.tminsum[v], 2 * (uint64_t)t.p2);
    return ts == -INFINITY ? -INFINITY : ub;
}

// Score of the object at partition-order position pos, terms in query order.
__device__ __forceinline__ float batch_score_one(const BatchQuery &Q, const float *__restrict__ pcols, uint64_t ld,
                                                 uint64_t pos) {
    float s = 0.0f;
    for (uint32_t j = 0; j < Q.m; j++) {
        const float pr = pcols[(uint64_t)Q.attr[j] * ld + pos] * Q.w[j];
        s = s + pr;
    }
    return s;
}

// Seed pass, one workgroup per query.  Wave 0 walks the bound tree from the
// root keeping the 4 children with the largest bounds at every level (a
// beam search); the first objects of the 4 partitions it reaches (16384 in
// all) are scored, and the exact k-th largest key among them is a lower
// bound of the query's k-th key: its score is tau.  Fewer than k objects:
// tau = -inf (every object is a candidate).
template <int NB>
__global__ __launch_bounds__(BSEED_BS) void batch_seed_kernel(BatchQuery *__restrict__ qs,
                                                              const BoundArgs *__restrict__ qb, TreeArgs t,
                                                              uint32_t L, const float *__restrict__ pcols,
                                                              uint64_t ld, const uint32_t *__restrict__ poff,
                                                              const uint32_t *__restrict__ psz,
                                                              const uint32_t *__restrict__ pid,
                                                              const uint32_t *__restrict__ qflag, int marked) {
    __shared__ SelectScratch ss;
    __shared__ uint32_t best[BSEED_T];
    __shared__ uint32_t nbest;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    if (qflag && qflag[q] != QF_SEED) return;  // second pass: only queries that need a seed
    if (marked && !qs[q].seed) return;         // learned first pass: only the queries the predictor marked
    const BatchQuery &Q = qs[q];
    if (tid < 64) {
        const BoundArgs &b = qb[q];
        const uint32_t lane = tid;
        uint32_t beam[BSEED_T] = {1u, 0u, 0u, 0u};
        uint32_t nb = 1;
        while (beam[0] < t.p2) {  // all beam nodes sit on one level
            uint64_t key = 0;
            if (lane < 2 * nb) {
                uint32_t parent = beam[0];
#pragma unroll
                for (uint32_t i = 1; i < BSEED_T; i++) parent = lane / 2 == i ? beam[i] : parent;
                const uint32_t c = 2 * parent + (lane & 1);
                const float ub = node_bound<NB>(b, t, c);
                if (ub > -INFINITY) key = ((uint64_t)okey(ub) << 32) | c;
            }
            uint32_t got = 0;
#pragma unroll
            for (uint32_t r = 0; r < BSEED_T; r++) {
                const uint64_t m = wave_max_u64(key);
                if (m == 0) break;
                beam[r] = (uint32_t)(m & 0xFFFFFFFFu);
                key = key == m ? 0ull : key;
                got++;
            }
            nb = got;
            if (!nb) break;
        }
        if (lane == 0) {
            uint32_t c = 0;
            for (uint32_t i = 0; i < nb; i++) {
                const uint32_t p = beam[i] - t.p2;
                if (beam[i] >= t.p2 && p < L && psz[p]) best[c++] = p;
            }
            nbest = c;
        }
    }
    __syncthreads();
    const uint32_t T = nbest;
    const uint32_t per = T ? BSEED_OBJS / T : 0u;
    uint32_t start[BSEED_T], lim[BSEED_T], total = 0;
#pragma unroll
    for (uint32_t i = 0; i < BSEED_T; i++) {
        const uint32_t p = i < T ? best[i] : 0u;
        start[i] = i < T ? poff[p] : 0u;
        lim[i] = i < T ? (psz[p] < per ? psz[p] : per) : 0u;
        total += lim[i];
    }
    // object slot j of this thread: position pos[j] (valid when ok[j]); the
    // terms run outside the objects so each term's loads are in flight together
    uint32_t pos[BSEED_KPT];
    bool ok[BSEED_KPT];
    float sc[BSEED_KPT];
#pragma unroll
    for (int j = 0; j < BSEED_KPT; j++) {
        const uint32_t g = j * BSEED_BS + tid;
        const uint32_t i = per ? g / per : BSEED_T, o = per ? g - i * per : 0u;
        uint32_t st = 0, lm = 0;
#pragma unroll
        for (uint32_t u = 0; u < BSEED_T; u++) {
            st = i == u ? start[u] : st;
            lm = i == u ? lim[u] : lm;
        }
        ok[j] = i < T && o < lm;
        pos[j] = ok[j] ? st + o : 0u;
        sc[j] = 0.0f;
    }
    for (uint32_t t = 0; t < Q.m; t++) {
        const float *col = pcols + (uint64_t)Q.attr[t] * ld;
        const float w = Q.w[t];
#pragma unroll
        for (int j = 0; j < BSEED_KPT; j++) {
            const float pr = col[pos[j]] * w;
            sc[j] = sc[j] + pr;
        }
    }
    RegKeys<BSEED_KPT> rk;
#pragma unroll
    for (int j = 0; j < BSEED_KPT; j++) rk.k[j] = ok[j] ? rkey(sc[j], pid[pos[j]]) : 0ull;
    float tau = -INFINITY;
    uint32_t tau_id = 0xFFFFFFFFu;
    if (total >= Q.k) {
        const uint64_t kth = block_kth_largest<BSEED_BS>(rk, total, Q.k, ss);
        tau = rkey_score(kth);
        tau_id = rkey_id(kth);
    }
    if (tid == 0) {
        qs[q].tau = tau;
        qs[q].tau_id = tau_id;
    }
}

// Tie scan, learned first pass: one workgroup per query the predictor
// marked as near its largest score (seed = 1).  An object whose every
// queried attribute of nonzero weight holds the store's extreme value on
// the side its weight prefers (largest for w > 0, smallest for w < 0) scores
// exactly smax: its terms are smax's terms.  Such objects are found by ANDing
// the extreme-value bitmaps of the query's attributes, in id order; if at
// least k exist, the k-th of them in id order has an id no smaller than the
// k-th smallest id among all objects scoring smax (they are a subset), so
// tau = (smax, that id) is a key no higher than the query's k-th key -- and
// equal to it when no other object rounds to smax.  The mark is then
// cleared.  Each step ANDs W consecutive words per thread (BXB_BS x 64 x W
// ids per step) and a block scan of the per-thread counts places the k-th.
// The scan gives up past `budget` ids, or once the ties counted so far
// project more than `budget` ids to reach k; such queries then take all
// ties as candidates or their learned score threshold, by the projected
// number of ties (end of the kernel).  No query leaves it marked.
constexpr uint64_t BXB_BUDGET = 1ull << 20;  // ids a scan may cover
constexpr int BXB_BS = 256;                  // threads per scan workgroup
constexpr int BXB_R0 = 4;                    // rows ANDed for every word
// W words of one row from 16-byte loads (rows are 128-byte aligned and w
// is a multiple of W).
template <int W>
__device__ __forceinline__ void load_words(const uint64_t *__restrict__ p, uint64_t (&x)[W]) {
    static_assert(W % 2 == 0, "words per thread");
#pragma unroll
    for (int v = 0; v < W / 2; v++) {
        const ulonglong2 t = reinterpret_cast<const ulonglong2 *>(p)[v];
        x[2 * v] = t.x;
        x[2 * v + 1] = t.y;
    }
}
// MAXR bounds the query's rows (terms of nonzero weight; 16 or 32).  Per
// step a thread ANDs the first BXB_R0 rows of its W words, then reads the
// other rows only if that partial AND is nonzero -- past a few attributes
// most words hold no common tie, and their remaining rows are never
// fetched.  The rows of each phase are loaded all at once.
template <int MAXR, int W>
__global__ __launch_bounds__(BXB_BS) void batch_tiebits_kernel(BatchQuery *__restrict__ qs,
                                                            const uint64_t *__restrict__ xbits, uint32_t d,
                                                            uint64_t nw, uint64_t n, uint64_t budget) {
    __shared__ uint32_t wsum[BXB_BS / 64], wtot[2][BXB_BS / 64];
    __shared__ uint32_t total, nrows;
    __shared__ uint32_t rrow[MAXR];
    const uint32_t tid = threadIdx.x, q = blockIdx.x;
    if (!qs[q].seed) return;  // uniform over the workgroup
    const BatchQuery &Q = qs[q];
    const uint32_t k = Q.k, m = Q.m;
    if (tid < 64) {
        // lane t: term t; the terms of nonzero weight, in term order, name
        // the rows (the weight's sign picks the maximum or the minimum row)
        const float w = tid < m ? Q.w[tid] : 0.0f;
        const uint32_t at = tid < m ? Q.attr[tid] : 0u;
        const uint64_t nz = __ballot(w != 0.0f);
        const uint32_t c = (uint32_t)__popcll(nz & ((1ull << tid) - 1ull));
        if (w != 0.0f && c < (uint32_t)MAXR) rrow[c] = w > 0.0f ? at : d + at;
        const uint32_t cnt = (uint32_t)__popcll(nz);
        const uint32_t nrw = cnt < (uint32_t)MAXR ? cnt : (uint32_t)MAXR;
        if (tid >= nrw && tid < (uint32_t)MAXR) rrow[tid] = 0u;  // unused rows: never loaded
        if (tid == 0) nrows = nrw;
    }
    __syncthreads();
    const uint32_t nr = __builtin_amdgcn_readfirstlane(nrows);
    const uint64_t *rowp[MAXR];  // uniform: scalar registers
#pragma unroll
    for (int r = 0; r < MAXR; r++) rowp[r] = xbits + (uint64_t)__builtin_amdgcn_readfirstlane(rrow[r]) * nw;
    const uint64_t lim = n < budget ? n : budget, wlim = (lim + 63) / 64;
    uint32_t found = 0, step = 0;
    uint64_t scanned = lim;  // ids read when the scan stops
    for (uint64_t w0 = 0; w0 < wlim; w0 += BXB_BS * W) {
        const uint64_t wb = w0 + (uint64_t)tid * W;
        // a thread past wlim reads words 0.. (in bounds) and masks them out;
        // words in [wlim, wb + W) lie inside the padded row
        const uint64_t wr = wb < wlim ? wb : 0;
        uint64_t acc[W];
#pragma unroll
        for (int i = 0; i < W; i++) acc[i] = wb + i < wlim ? ~0ull : 0ull;
        {
            uint64_t x[BXB_R0][W];
#pragma unroll
            for (int r = 0; r < BXB_R0; r++) {
                if ((uint32_t)r < nr) {
                    load_words<W>(rowp[r] + wr, x[r]);
                } else {
#pragma unroll
                    for (int i = 0; i < W; i++) x[r][i] = ~0ull;
                }
            }
#pragma unroll
            for (int r = 0; r < BXB_R0; r++)
#pragma unroll
                for (int i = 0; i < W; i++) acc[i] &= x[r][i];
        }
        uint64_t any = 0;
#pragma unroll
        for (int i = 0; i < W; i++) any |= acc[i];
        if (nr > (uint32_t)BXB_R0 && any) {
            uint64_t x[MAXR - BXB_R0][W];
#pragma unroll
            for (int r = BXB_R0; r < MAXR; r++) {
                if ((uint32_t)r < nr) {
                    load_words<W>(rowp[r] + wr, x[r - BXB_R0]);
                } else {
#pragma unroll
                    for (int i = 0; i < W; i++) x[r - BXB_R0][i] = ~0ull;
                }
            }
#pragma unroll
            for (int r = 0; r < MAXR - BXB_R0; r++)
#pragma unroll
                for (int i = 0; i < W; i++) acc[i] &= x[r][i];
        }
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (wb + i == wlim - 1 && lim % 64) acc[i] &= (1ull << (lim % 64)) - 1ull;  // the last word may run past lim
            c += (uint32_t)__popcll(acc[i]);
        }
        // the step's count: wave sums through a double-buffered LDS row, one
        // barrier; the prefix scan runs only in the step that reaches k
        const uint32_t cw = wave_sum_u32(c);
        if (lane_id() == 0) wtot[step & 1][tid / 64] = cw;
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int i = 0; i < BXB_BS / 64; i++) tot += wtot[step & 1][i];
        step++;
        if (found + tot >= k) {  // uniform
            const uint32_t excl = block_excl_scan<BXB_BS>(c, wsum, &total);
            if (found + excl < k && k <= found + excl + c) {  // this thread holds the k-th
                uint32_t r = k - found - excl - 1;
#pragma unroll
                for (int i = 0; i < W; i++) {
                    const uint32_t pc = (uint32_t)__popcll(acc[i]);
                    if (r < pc) {
                        uint64_t y = acc[i];
                        for (uint32_t j = 0; j < r; j++) y &= y - 1ull;
                        BatchQuery &Wq = qs[q];
                        Wq.tau = Q.smax;
                        Wq.tau_id = (uint32_t)(64 * (wb + i)) + (uint32_t)(__ffsll((long long)y) - 1);
                        Wq.seed = 0u;
                        break;
                    }
                    r -= pc;
                }
            }
            return;
        }
        found += tot;
        const uint64_t done = (w0 + BXB_BS * W) * 64;
        if ((uint64_t)found * budget < (uint64_t)k * done) {  // projected past the budget
            scanned = done;
            break;
        }
    }
    // The scan stopped short of k ties; the rate read projects the store's.
    // At least 2k of them: tau = (smax, every id) takes them all as
    // candidates (a key no higher than the k-th when at least k exist).
    // Fewer: the predictor's mark was most likely wrong (about k objects or
    // fewer reach smax), so the query keeps its learned score threshold.
    // Either is exact: a threshold that leaves fewer than k candidates, or
    // more than the list holds, sends the query to the second pass.
    if (tid == 0) {
        const uint64_t proj = (uint64_t)found * n;  // ties in the store x ids read
        if (proj >= 2ull * k * scanned) {
            BatchQuery &Wq = qs[q];
            Wq.tau = Q.smax;
            Wq.tau_id = 0xFFFFFFFFu;
        }
        qs[q].seed = 0u;
    }
}

// Query-major mask: workgroup (q, y) of S per query walks the bound tree
// two levels deep.  Groups y, y + S, ... of 2^H partitions (tree nodes
// G + g) whose bound reaches tau_q go to an LDS list; every partition of a
// listed group whose own bound reaches tau_q gets q appended to its query
// list plist[p * nq ...] (one atomic per (partition, query) pair on the
// partition's counter).  Each thread evaluates 32 / NB bounds at a time so
// their loads and atomics overlap.
template <int NB>
__global__ __launch_bounds__(BMASK_BS) void batch_mask_query_kernel(const BatchQuery *__restrict__ qs,
                                                                    const BoundArgs *__restrict__ qb, TreeArgs t,
                                                                    uint32_t G, uint32_t H, uint32_t L, uint32_t nq,
                                                                    unsigned int *__restrict__ pcnt,
                                                                    uint16_t *__restrict__ plist,
                                                                    const uint32_t *__restrict__ qflag) {
    constexpr int U = 32 / NB;
    __shared__ uint32_t list[BMASK_PAIRS];
    __shared__ uint32_t nlist;
    const uint32_t q = blockIdx.x, y = blockIdx.y, S = gridDim.y, tid = threadIdx.x;
    if (qflag && qflag[q] == QF_DONE) return;  // second pass: settled queries take no part
    if (tid == 0) nlist = 0;
    __syncthreads();
    const BoundArgs &b = qb[q];
    const float tau = qs[q].tau;
    const bool sat = at_smax(qs[q]);
    const uint32_t tau_id = qs[q].tau_id;
    for (uint32_t g = y + S * tid; g < G; g += S * BMASK_BS) {
        const float ub = node_bound<NB>(b, t, G + g);
        if (ub > -INFINITY && ub >= tau && !(sat && t.tminid[G + g] > tau_id)) list[atomicAdd(&nlist, 1u)] = g;
    }
    __syncthreads();
    const uint32_t nl = nlist, per = 1u << H, total = nl * per;
    for (uint32_t i0 = tid; i0 < total; i0 += U * BMASK_BS) {
        uint32_t p[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * BMASK_BS;
            p[u] = i < total ? (list[i >> H] << H) + (i & (per - 1)) : L;
            const float ub = p[u] < L ? node_bound<NB>(b, t, t.p2 + p[u]) : -INFINITY;
            in[u] = ub > -INFINITY && ub >= tau && !(sat && t.tminid[t.p2 + p[u]] > tau_id);
        }
        uint32_t pos[U];
#pragma unroll
        for (int u = 0; u < U; u++) pos[u] = in[u] ? atomicAdd(&pcnt[p[u]], 1u) : 0u;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u]) plist[(uint64_t)p[u] * nq + pos[u]] = (uint16_t)q;
    }
}

// Group-major mask: workgroup (g, y) takes group g of 2^H partitions (tree node G + g) and
// the query slice [y * QS, (y + 1) * QS): the queries whose bound over the
// group reaches tau_q go to an LDS list, then every (listed query,
// partition of the group) pair whose bound reaches tau_q is counted per
// partition in LDS.  One global atomic per partition reserves room in its
// query list plist[p * nq ...], so a partition that many queries reach sees
// one atomic per workgroup, not one per query.  Each thread evaluates
// 32 / NB pair bounds at a time so their loads overlap.
template <int NB>
__global__ __launch_bounds__(BMASK_BS) void batch_mask_kernel(const BatchQuery *__restrict__ qs,
                                                              const BoundArgs *__restrict__ qb, TreeArgs t,
                                                              uint32_t G, uint32_t H, uint32_t L, uint32_t nq,
                                                              uint32_t QS, unsigned int *__restrict__ pcnt,
                                                              uint16_t *__restrict__ plist,
                                                              const uint32_t *__restrict__ qflag) {
    constexpr int U = 32 / NB;  // bounds in flight per thread (registers: ~4 NB per bound)
    __shared__ uint16_t lq[BMASK_PAIRS];   // listed queries
    __shared__ uint16_t ent[BMASK_PAIRS];  // pair (i, j): 1 + its slot in partition j's share, 0 = out
    __shared__ uint32_t lcnt[BMASK_MAX_PER], lbase[BMASK_MAX_PER];
    __shared__ uint32_t nlist;
    const uint32_t g = blockIdx.x, tid = threadIdx.x, per = 1u << H;
    if (tid == 0) nlist = 0;
    if (tid < per) lcnt[tid] = 0;
    __syncthreads();
    const uint32_t qa = blockIdx.y * QS, qe = nq - qa < QS ? nq : qa + QS;
    for (uint32_t q = qa + tid; q < qe; q += BMASK_BS) {
        if (qflag && qflag[q] == QF_DONE) continue;  // second pass: settled queries take no part
        const float ub = node_bound<NB>(qb[q], t, G + g);
        if (ub > -INFINITY && ub >= qs[q].tau && !(at_smax(qs[q]) && t.tminid[G + g] > qs[q].tau_id))
            lq[atomicAdd(&nlist, 1u)] = (uint16_t)q;
    }
    __syncthreads();
    const uint32_t nl = nlist;
    if (nl == 0) return;
    const uint32_t total = nl << H, p0 = g << H;
    for (uint32_t i0 = tid; i0 < total; i0 += U * BMASK_BS) {
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * BMASK_BS, p = p0 + (i & (per - 1));
            in[u] = false;
            if (i < total && p < L) {
                const uint32_t q = lq[i >> H];
                const float ub = node_bound<NB>(qb[q], t, t.p2 + p);
                in[u] = ub > -INFINITY && ub >= qs[q].tau && !(at_smax(qs[q]) && t.tminid[t.p2 + p] > qs[q].tau_id);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * BMASK_BS;
            if (i < total) ent[i] = in[u] ? (uint16_t)(atomicAdd(&lcnt[i & (per - 1)], 1u) + 1u) : (uint16_t)0;
        }
    }
    __syncthreads();
    if (tid < per) {
        const uint32_t c = lcnt[tid];
        lbase[tid] = c ? atomicAdd(&pcnt[p0 + tid], c) : 0u;
    }
    __syncthreads();
    for (uint32_t i = tid; i < total; i += BMASK_BS) {
        const uint32_t e = ent[i];
        if (e) {
            const uint32_t j = i & (per - 1);
            plist[(uint64_t)(p0 + j) * nq + lbase[j] + e - 1] = lq[i >> H];
        }
    }
}

// One thread per partition: its query list is cut into ranges of at most
// qmax queries and every range into object chunks of ochunk.  Ranges holding
// at least qmax / 2 queries are "heavy" and fill the item list from the
// front, the others from the back (index cap - 1 - i), so the persistent
// scoring kernel starts the long items first; only a list's last range can
// be light.  One global atomic per workgroup reserves its items, two more
// add its pair counts.  The items of a partition are written by the whole
// wave (a partition many queries reach has a hundred or more).
__global__ __launch_bounds__(BAUX_BS) void batch_compact_kernel(const unsigned int *__restrict__ pcnt, uint32_t L,
                                                                uint32_t qmax, uint32_t ochunk, uint64_t cap,
                                                                const uint32_t *__restrict__ psz,
                                                                BatchItem *__restrict__ items,
                                                                unsigned int *__restrict__ ctr) {
    __shared__ uint32_t wsum[BAUX_BS / 64];
    __shared__ unsigned long long wstat[BAUX_BS / 64][2];
    __shared__ uint32_t tot_h, tot_l, base_h, base_l;
    const uint32_t p = blockIdx.x * BAUX_BS + threadIdx.x, lane = lane_id();
    const uint32_t cnt = p < L ? pcnt[p] : 0u;
    const uint32_t sz = cnt ? psz[p] : 0u;
    const uint32_t nch = (sz + ochunk - 1) / ochunk;
    const uint32_t heavy_min = (qmax + 1) / 2;
    const uint32_t nr = (cnt + qmax - 1) / qmax;
    const uint32_t last = cnt - (nr ? (nr - 1) * qmax : 0u);
    const uint32_t nhr = nr ? nr - 1 + (last >= heavy_min ? 1u : 0u) : 0u;
    const uint32_t nh = nhr * nch, nl = (nr - nhr) * nch;
    const uint32_t offh = block_excl_scan<BAUX_BS>(nh, wsum, &tot_h);
    const uint32_t offl = block_excl_scan<BAUX_BS>(nl, wsum, &tot_l);
    unsigned long long pairs = cnt, objs = (unsigned long long)cnt * sz;
    for (int o = 32; o > 0; o >>= 1) {
        pairs += __shfl_xor(pairs, o, 64);
        objs += __shfl_xor(objs, o, 64);
    }
    if (lane == 0) {
        wstat[threadIdx.x / 64][0] = pairs;
        wstat[threadIdx.x / 64][1] = objs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // heavy items count in ctr[0], light in ctr[1]: one 64-bit atomic
        // reserves both (the low word never carries)
        const unsigned long long r =
            tot_h | tot_l ? atomicAdd(reinterpret_cast<unsigned long long *>(ctr),
                                      ((unsigned long long)tot_l << 32) | tot_h)
                          : 0ull;
        base_h = (uint32_t)r;
        base_l = (uint32_t)(r >> 32);
        unsigned long long sp = 0, so = 0;
        for (int w = 0; w < BAUX_BS / 64; w++) {
            sp += wstat[w][0];
            so += wstat[w][1];
        }
        if (sp) {
            atomicAdd(reinterpret_cast<unsigned long long *>(ctr + 2), sp);
            atomicAdd(reinterpret_cast<unsigned long long *>(ctr + 4), so);
        }
    }
    __syncthreads();
    const uint32_t kh0 = base_h + offh, kl0 = base_l + offl;
    for (uint64_t m = __ballot(nr != 0); m; m &= m - 1) {
        const int src = __ffsll((long long)m) - 1;
        const uint32_t sp_ = __shfl(p, src, 64), scnt = __shfl(cnt, src, 64), snch = __shfl(nch, src, 64);
        const uint32_t snr = __shfl(nr, src, 64), snhr = __shfl(nhr, src, 64);
        const uint32_t skh = __shfl(kh0, src, 64), skl = __shfl(kl0, src, 64);
        for (uint32_t i = lane; i < snr * snch; i += 64) {
            const uint32_t r = i / snch, c = i - r * snch;
            const uint32_t q0 = r * qmax, qn = scnt - q0 < qmax ? scnt - q0 : qmax;
            const BatchItem it{sp_, q0, qn, c};
            if (r < snhr) items[skh + i] = it;           // heavy ranges come first
            else items[cap - 1 - (skl + c)] = it;        // the light last range
        }
    }
}

struct ScoreArgs {
    const float *pcols;
    uint64_t ld;
    uint32_t d;
    const uint32_t *poff, *psz, *pid;
    const BatchItem *items;
    uint64_t cap;       // item list capacity (light items sit at its end)
    uint32_t ochunk;    // objects per item
    const uint16_t *plist;
    uint32_t nq;
    unsigned int *ctr;
    const BatchQuery *q;
    uint64_t *cand;
    unsigned int *ccount;
    const float *cstat;          // chunk statistics, element e of chunk c at [e * nchunks + c]
    uint32_t crow;               // elements per chunk
    uint64_t nchunks;
    unsigned long long *scored;  // per workgroup: (object, query) pairs scored
    const uint32_t *cminid;      // smallest object id per chunk
};

template <int VPT>
__device__ __forceinline__ void load_objs(const float *__restrict__ p, float (&x)[VPT]) {
    if constexpr (VPT % 4 == 0) {
#pragma unroll
        for (int v = 0; v < VPT / 4; v++) {
            const f4 t = *reinterpret_cast<const f4 *>(p + 4 * v);
            x[4 * v] = t[0];
            x[4 * v + 1] = t[1];
            x[4 * v + 2] = t[2];
            x[4 * v + 3] = t[3];
        }
    } else {
        static_assert(VPT == 2, "objects per thread");
        const float2 t = *reinterpret_cast<const float2 *>(p);
        x[0] = t.x;
        x[1] = t.y;
    }
}

#define TK_TERM(c)                                       \
    case c:                                              \
        if constexpr ((c) < D) {                         \
            _Pragma("unroll") for (int e = 0; e < VPT; e++) { \
                const float pr = x[c][e] * w;            \
                s[e] = s[e] + pr;                        \
            }                                            \
        }                                                \
        break;

// Persistent scoring kernel: D = attribute registers per object (>= d),
// VPT consecutive objects per thread, BS threads per workgroup.
//
// Latency: the query list of an item and, for D <= 16, every listed query's
// ascending-form weights and threshold are loaded once per item, one query
// per lane; the per-query loop reads them with readlane, so it waits on no
// memory.  The ticket of the next item is taken after the current item's
// loads are issued and is read only when the item is done.
//
// Chunk pruning: objects are sorted by attribute sum inside a partition and
// the statistics row of a chunk of PTA_CHUNK positions bounds that chunk and
// every later one of the partition (suffix statistics).  Lane j evaluates
// the LP bound of query j from the row of the wave's first chunk (ascending
// queries with D <= 16; other queries always pass); the wave scores only the
// queries whose bound reaches their threshold, loads its objects only if one
// does, and a query that fails stays out of the wave's later tiles, which
// ends the wave's walk through the item once every query has failed.
template <int D, int VPT, int BS>
__global__ __launch_bounds__(BS) void batch_score_kernel(ScoreArgs a) {
    constexpr bool LANE_W = D <= 16;
    __shared__ uint32_t s_it[2];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t nheavy = __builtin_amdgcn_readfirstlane(a.ctr[0]);
    const uint32_t nlight = __builtin_amdgcn_readfirstlane(a.ctr[1]);
    const uint32_t nitems = nheavy + nlight;
    if (tid == 0) s_it[0] = atomicAdd(&a.ctr[6], 1u);
    __syncthreads();
    uint32_t it = __builtin_amdgcn_readfirstlane(s_it[0]);
    uint32_t slot = 1;
    uint32_t npairs = 0;
    unsigned long long nterms = 0;
    while (it < nitems) {
        const BatchItem I = it < nheavy ? a.items[it] : a.items[a.cap - 1 - (it - nheavy)];
        const uint32_t first = I.chunk * a.ochunk, psz_ = a.psz[I.p];
        const uint32_t start = a.poff[I.p] + first;
        const uint32_t size = psz_ - first < a.ochunk ? psz_ - first : a.ochunk;
        const uint16_t *ql = a.plist + (uint64_t)I.p * a.nq + I.qbeg;
        uint32_t nxt = 0;
        bool ticket = tid == 0;
        for (uint32_t q0 = 0; q0 < I.qcnt; q0 += 64) {
            const uint32_t nqb = I.qcnt - q0 < 64u ? I.qcnt - q0 : 64u;
            // lane j: query q0 + j of the item (lanes past nqb hold query 0)
            const uint32_t lq = lane < nqb ? (uint32_t)ql[q0 + lane] : 0u;
            const f4 lt = *reinterpret_cast<const f4 *>(&a.q[lq].tau);  // tau, tau_id, ascending, m
            const bool lsat = at_smax(a.q[lq]);  // lane j's query sits at its largest score
            float lw[LANE_W ? D : 1];
            f4 lws[LANE_W ? D / 4 : 1];
            uint4 las = {0u, 0u, 0u, 0u};
            if constexpr (LANE_W) {
#pragma unroll
                for (int c = 0; c < D; c += 4) {
                    const f4 v = *reinterpret_cast<const f4 *>(&a.q[lq].wa[c]);
                    lw[c] = v[0];
                    lw[c + 1] = v[1];
                    lw[c + 2] = v[2];
                    lw[c + 3] = v[3];
                    lws[c / 4] = reinterpret_cast<const f4 *>(a.q[lq].ws)[c / 4];
                }
                las = *reinterpret_cast<const uint4 *>(a.q[lq].as);
            }
            if (ticket) {
                nxt = atomicAdd(&a.ctr[6], 1u);
                ticket = false;
            }
            // lanes whose query may still reach its threshold on the wave's
            // remaining objects (suffix bounds: once below, below for good)
            bool alive = lane < nqb;
            for (uint32_t b0 = 0; b0 < size; b0 += BS * VPT) {
                const uint32_t i0 = b0 + tid * VPT;
                const uint32_t nvalid = i0 < size ? (size - i0 < (uint32_t)VPT ? size - i0 : (uint32_t)VPT) : 0u;
                // this wave's objects: item positions [wb, ...)
                const uint32_t wb = __builtin_amdgcn_readfirstlane(b0 + (tid & ~63u) * VPT);
                if (wb >= size || !__ballot(alive)) break;
                // suffix bound of lane j's query from the row of the wave's
                // first chunk, evaluated by every lane (the shuffles read all
                // lanes)
                bool pass = alive;
                if constexpr (LANE_W) {
                    // the lane's weights and attribute bytes stay in this
                    // compact form across tiles: the bound's FP64 weights and
                    // shuffle addresses are rebuilt per tile, not kept live
                    // (at D = 16: 154 registers instead of 199, three waves
                    // per SIMD instead of two)
#pragma unroll
                    for (int i = 0; i < D / 4; i++) asm volatile("" : "+v"(lws[i]));
                    asm volatile("" : "+v"(las.x), "+v"(las.y), "+v"(las.z), "+v"(las.w));
                    const uint32_t dd = a.d;
                    const uint32_t c0 = (start + wb) / PTA_CHUNK;
                    const float cs = lane < a.crow ? a.cstat[(uint64_t)lane * a.nchunks + c0] : 0.0f;  // lane e: element e
                    const float csum = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs), (int)(2 * dd)));
                    const float cminsum =
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs), (int)(2 * dd + 1)));
                    const uint32_t mq = (uint32_t)__float_as_int(lt[3]);
                    double wv[D];
                    float mxv[D], mnv[D];
#pragma unroll
                    for (int i = 0; i < D; i += 4) {
                        const f4 w4 = lws[i / 4];
                        const uint32_t aw = i / 4 == 0 ? las.x : (i / 4 == 1 ? las.y : (i / 4 == 2 ? las.z : las.w));
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const bool on = (uint32_t)(i + u) < mq;
                            const int at = on ? (int)((aw >> (8 * u)) & 0xFFu) : 0;
                            const float mx = __shfl(cs, at, 64), mn = __shfl(cs, (int)dd + at, 64);
                            wv[i + u] = on ? (double)w4[u] : 0.0;
                            mxv[i + u] = on ? mx : 0.0f;
                            mnv[i + u] = on ? mn : 0.0f;
                        }
                    }
                    const float ub = lp_bound_regs<D>(wv, mxv, mnv, csum, cminsum);
                    const bool asc_l = __float_as_int(lt[2]) != 0;
                    pass = pass && (!asc_l || ub >= lt[0]);
                    alive = pass;  // a failing suffix bound covers every later object of the partition
                }
                // queries at their largest score skip tiles whose ids all
                // exceed tau_id (not a suffix property: alive is unchanged)
                if (__ballot(pass && lsat)) {
                    constexpr uint32_t TILE = 64 * VPT;
                    const uint32_t c0 = (start + wb) / PTA_CHUNK;
                    const uint32_t cn = (TILE + PTA_CHUNK - 1) / PTA_CHUNK;
                    uint32_t wmin = 0xFFFFFFFFu;
                    for (uint32_t i = 0; i < cn && c0 + i < a.nchunks; i++) {  // may reach the next partition: still a bound
                        const uint32_t mi = a.cminid[c0 + i];
                        wmin = mi < wmin ? mi : wmin;
                    }
                    pass = pass && !(lsat && wmin > __float_as_uint(lt[1]));
                }
                const uint64_t qmask = __ballot(pass);
                if (!qmask) continue;
                // the objects, loaded once the wave has queries to score
                float x[D][VPT];
#pragma unroll
                for (int c = 0; c < D; c++) {
                    if (c < (int)a.d && nvalid) {
                        load_objs<VPT>(a.pcols + (uint64_t)c * a.ld + start + i0, x[c]);
                    } else {
#pragma unroll
                        for (int e = 0; e < VPT; e++) x[c][e] = 0.0f;
                    }
                }
                for (uint64_t mm = qmask; mm; mm &= mm - 1) {
                    const uint32_t qi = (uint32_t)__ffsll((long long)mm) - 1u;
                    npairs += nvalid;
                    nterms += (unsigned long long)nvalid *
                              (uint32_t)__builtin_amdgcn_readlane(__float_as_int(lt[3]), (int)qi);
                    const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)lq, (int)qi);
                    const uint32_t asc = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(lt[2]), (int)qi);
                        float s[VPT];
#pragma unroll
                        for (int e = 0; e < VPT; e++) s[e] = 0.0f;
                        if (asc) {
                            // static register order, zero weights skipped by a uniform branch
#pragma unroll
                            for (int c = 0; c < D; c++) {
                                float w;
                                if constexpr (LANE_W)
                                    w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lw[LANE_W ? c : 0]),
                                                                                 (int)qi));
                                else
                                    w = a.q[q].wa[c];
                                if (w != 0.0f) {
#pragma unroll
                                    for (int e = 0; e < VPT; e++) {
                                        const float pr = x[c][e] * w;
                                        s[e] = s[e] + pr;
                                    }
                                }
                            }
                        } else {
                            const BatchQuery &Q = a.q[q];
                            for (uint32_t j = 0; j < Q.m; j++) {
                                const float w = Q.w[j];
                                switch (Q.attr[j]) {
                                    TK_TERM(0) TK_TERM(1) TK_TERM(2) TK_TERM(3) TK_TERM(4) TK_TERM(5) TK_TERM(6)
                                    TK_TERM(7) TK_TERM(8) TK_TERM(9) TK_TERM(10) TK_TERM(11) TK_TERM(12) TK_TERM(13)
                                    TK_TERM(14) TK_TERM(15) TK_TERM(16) TK_TERM(17) TK_TERM(18) TK_TERM(19) TK_TERM(20)
                                    TK_TERM(21) TK_TERM(22) TK_TERM(23) TK_TERM(24) TK_TERM(25) TK_TERM(26) TK_TERM(27)
                                    TK_TERM(28) TK_TERM(29) TK_TERM(30) TK_TERM(31)
                                    default: break;
                                }
                            }
                        }
                        // admitted: key (s, id) >= (tau, tau_id); the id is read only on a tie
                        const float tau = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lt[0]), (int)qi));
                        const uint32_t tau_id = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(lt[1]), (int)qi);
                        uint32_t adm = 0;
#pragma unroll
                        for (int e = 0; e < VPT; e++) {
                            bool in = (uint32_t)e < nvalid && s[e] >= tau;
                            if (in && s[e] == tau && tau_id != 0xFFFFFFFFu) in = a.pid[start + i0 + e] <= tau_id;
                            adm |= in ? 1u << e : 0u;
                        }
                        const uint32_t np = (uint32_t)__popc(adm);
                        if (__ballot(np != 0)) {
                            const uint32_t incl = wave_incl_scan(np);
                            const uint32_t tot = __shfl(incl, 63, 64);
                            uint32_t base = 0;
                            if (lane == 63) base = atomicAdd(&a.ccount[q], tot);
                            base = __shfl(base, 63, 64);
                            uint32_t pos = base + incl - np;
                            uint64_t *out = a.cand + (uint64_t)q * BATCH_CAP;
#pragma unroll
                            for (int e = 0; e < VPT; e++) {
                                if ((adm >> e) & 1u) {
                                    if (pos < BATCH_CAP) out[pos] = rkey(s[e], a.pid[start + i0 + e]);
                                    pos++;
                                }
                            }
                        }
                    }
                }
            }
        if (ticket) nxt = atomicAdd(&a.ctr[6], 1u);  // no wave of tid 0 reached the queries
        if (tid == 0) s_it[slot] = nxt;
        __syncthreads();
        it = __builtin_amdgcn_readfirstlane(s_it[slot]);
        slot ^= 1u;
    }
    // scored (object, query) pairs and terms of this workgroup
    __shared__ unsigned long long wpairs[BS / 64][2];
    unsigned long long t = npairs, tt = nterms;
    for (int o = 32; o > 0; o >>= 1) {
        t += __shfl_xor(t, o, 64);
        tt += __shfl_xor(tt, o, 64);
    }
    if (lane == 0) {
        wpairs[tid / 64][0] = t;
        wpairs[tid / 64][1] = tt;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < BS / 64; w++) {
            t += wpairs[w][0];
            tt += wpairs[w][1];
        }
        a.scored[blockIdx.x] = t;
        a.scored[BATCH_MAX_GRID + blockIdx.x] = tt;
    }
}
#undef TK_TERM

struct FinalArgs {
    BatchQuery *q;
    const uint64_t *cand;
    unsigned int *ccount;    // first pass: read, then cleared for the second
    unsigned int *ctr;       // first pass: copied out, then cleared for the second
    uint32_t kstride;
    uint64_t *ids;
    float *scores;
    float *thr;
    uint32_t *status;
    uint32_t *cnt;
    uint32_t *ctr_out;
    uint32_t *qflag;   // first pass: written; second pass: read (settled queries are skipped)
    int second;
    const unsigned long long *scored;  // per scoring workgroup of this pass
    uint32_t nscored;
    int seed2;               // the second pass runs a seed threshold pass
    unsigned int *pcnt;      // first pass: cleared here for the second
    uint32_t L;
    unsigned int *flist;     // first pass: queries flagged for the second ([0] count, then ids); may be null
};

// Where a batch's results go: the caller's own arrays when every query
// fills its stride and they are pinned host memory (direct), else the
// slot's pinned staging block.  Staged with the queries, so a captured
// graph reads it from device memory.
struct OutDesc {
    uint64_t ids, scores, thr;
    uint32_t direct, pad;
};

// Publishes a finished batch to pinned host memory, as the last step of
// the batch on its slot's stream: direct, the ids, scores and thresholds
// into the caller's arrays and the status words into the staging block;
// otherwise the whole output block into the staging block.  Its time is
// the host link's, and the other slot's batch runs meanwhile.
constexpr uint32_t BPUB_GRID = 64;
struct PublishArgs {
    const unsigned char *d_out;
    size_t ids_off, sc_off, thr_off, st_off, bytes;
    uint64_t nres;
    uint32_t nq;
    const OutDesc *od;
    unsigned char *h_out;
};

__global__ __launch_bounds__(256) void batch_publish_kernel(PublishArgs a) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x, stride = (uint64_t)gridDim.x * 256;
    const uint32_t *dw = reinterpret_cast<const uint32_t *>(a.d_out);
    uint32_t *hw = reinterpret_cast<uint32_t *>(a.h_out);
    const OutDesc od = *a.od;
    if (od.direct) {
        const uint64_t *di = reinterpret_cast<const uint64_t *>(a.d_out + a.ids_off);
        const float *ds = reinterpret_cast<const float *>(a.d_out + a.sc_off);
        uint64_t *ids = reinterpret_cast<uint64_t *>(od.ids);
        float *scores = reinterpret_cast<float *>(od.scores), *thr = reinterpret_cast<float *>(od.thr);
        for (uint64_t i = t; i < a.nres; i += stride) {
            ids[i] = di[i];
            scores[i] = ds[i];
        }
        const float *dt = reinterpret_cast<const float *>(a.d_out + a.thr_off);
        for (uint64_t i = t; i < a.nq; i += stride) thr[i] = dt[i];
        for (uint64_t i = a.st_off / 4 + t; i < a.bytes / 4; i += stride) hw[i] = dw[i];
    } else {
        for (uint64_t i = t; i < a.bytes / 4; i += stride) hw[i] = dw[i];
    }
}

template <int DB>
__global__ __launch_bounds__(BFIN_BS) void batch_final_kernel(FinalArgs a) {
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ uint64_t sorted[TOPK_MAX_K];
    __shared__ SelectScratch ss;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    if (q == 0) {
        // counters of this pass; words 4..5 and 8..9 become the (object,
        // query) pairs and terms the scoring kernel actually scored
        if (tid < BCTR_WORDS) {
            if (tid < 4 || tid == 6 || tid == 7) a.ctr_out[tid] = a.ctr[tid];
            if (!a.second) a.ctr[tid] = 0u;  // counters of the second pass
        }
        unsigned long long t = 0, tt = 0;
        for (uint32_t i = tid; i < a.nscored; i += BFIN_BS) {
            t += a.scored[i];
            tt += a.scored[BATCH_MAX_GRID + i];
        }
        for (int o = 32; o > 0; o >>= 1) {
            t += __shfl_xor(t, o, 64);
            tt += __shfl_xor(tt, o, 64);
        }
        if (lane_id() == 0) {
            top[tid / 64] = t;
            sorted[tid / 64] = tt;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < BFIN_BS / 64; w++) {
                t += top[w];
                tt += sorted[w];
            }
            *reinterpret_cast<unsigned long long *>(a.ctr_out + 4) = t;
            *reinterpret_cast<unsigned long long *>(a.ctr_out + 8) = tt;
        }
        __syncthreads();
    }
    if (a.second && a.qflag[q] == QF_DONE) return;
    const uint32_t c = a.ccount[q], k = a.q[q].k;
    if (tid == 0) a.cnt[q] = c;
    if (!a.second) {
        // clear what the second pass counts: this query's candidates and a
        // share of the partition counters (the compaction has read them)
        __syncthreads();
        if (tid == 0) a.ccount[q] = 0u;
        for (uint64_t i = (uint64_t)q * BFIN_BS + tid; i < a.L; i += (uint64_t)gridDim.x * BFIN_BS) a.pcnt[i] = 0u;
    }
    if (c > BATCH_CAP) {
        if (a.second) {
            if (tid == 0) a.status[q] = STATUS_OVERFLOW;
            return;
        }
        // The stored keys are a subset of the objects, so their k-th largest
        // is a lower bound of the query's k-th key: the second pass reruns the
        // query with that tighter threshold.
        const uint64_t kth = block_kth_largest<BFIN_BS>(MemKeys<BFIN_BS>{a.cand + (uint64_t)q * BATCH_CAP, BATCH_CAP},
                                                        BATCH_CAP, k, ss);
        if (tid == 0) {
            a.status[q] = STATUS_OVERFLOW;
            a.q[q].tau = rkey_score(kth);
            a.q[q].tau_id = rkey_id(kth);
            a.qflag[q] = QF_GIVEN;
            if (a.flist) a.flist[1 + atomicAdd(&a.flist[0], 1u)] = q;
        }
        return;
    }
    if (c < k) {
        // a predicted threshold above the k-th score: the second pass reruns
        // the query with its lower retry threshold, or seeds it
        if (tid == 0) {
            a.status[q] = STATUS_RETRY;
            if (!a.second) {
                const float tr = a.q[q].tau_retry;
                if (tr < a.q[q].tau) {
                    a.q[q].tau = tr;
                    a.q[q].tau_id = 0xFFFFFFFFu;
                    a.qflag[q] = QF_LOWER;
                } else {
                    // no lower threshold: a seed pass, or (without one) the
                    // host's exact path (status RETRY stays)
                    a.qflag[q] = a.seed2 ? QF_SEED : QF_DONE;
                }
                if (a.flist && a.qflag[q] != QF_DONE) a.flist[1 + atomicAdd(&a.flist[0], 1u)] = q;
            }
        }
        return;
    }
    if (tid == 0 && !a.second) a.qflag[q] = QF_DONE;
    const FinalOut o{a.ids + (uint64_t)q * a.kstride, a.scores + (uint64_t)q * a.kstride, a.thr + q, a.status + q};
    select_and_emit<BFIN_BS, BFIN_KPT, FlatView, DB>(FlatView{a.cand + (uint64_t)q * BATCH_CAP}, c, k, o, top, sorted,
                                                     ss);
}

// ---------------------------------------------------------------------------
// query-major pass
// ---------------------------------------------------------------------------
// Three launches replace the mask, the compaction and the item scoring when
// thresholds are tight (a trained predictor): each query then reaches a few
// dozen partitions and a few chunks in each, and the time of the item path
// goes to list building, contended partition counters and dependent loads.
//
//  1. batch_cut_kernel, one workgroup per (query, slice s of S): the bound of
//     every group of 2^H partitions of the slice (groups s, s + S, ...);
//     the bounds of every partition of a group that reaches tau (BQ_U in
//     flight per thread); then, for the partitions that reach tau, the
//     suffix-row bound of every chunk (one thread per chunk over an
//     exclusive scan of their chunk counts): the first failing chunk cuts
//     the partition (every later chunk fails too).  The workgroup's unit
//     u = q * S + s lists (partition, first chunk index) pairs in
//     tasks[u * ucap ...], chunk indices as an exclusive prefix over the
//     list, and ucnt[u] = (entries, chunks).
//  2. batch_unit_scan_kernel (one workgroup): exclusive prefix of the units'
//     chunk counts, uoff[0..U].
//  3. batch_chunk_kernel: the uoff[U] chunks of all units are split into
//     equal contiguous ranges, one per wave; a wave walks its range (its
//     first unit and list entry found by a 64-way search per level), scores
//     each chunk (4 objects per lane, the query's own term order) and
//     appends the keys that reach (tau, tau_id) to the query's candidate
//     list (one atomic per wave and chunk on the query's counter).  A query
//     with many chunks is spread over many waves, so no query's chain sets
//     the pass's time.
constexpr int BQ_BS = 256;
constexpr uint32_t BQ_PCAP = 1024;       // listed partitions per flush (4 per thread in the scans)
constexpr uint32_t BQ_EPT = BQ_PCAP / BQ_BS;
constexpr uint32_t BQ_MAX_S = 64;        // slices per query
// A full batch of 1024 queries keeps every workgroup resident at 4 per CU
// when a wave holds at most 128 VGPRs (4 waves per SIMD).
constexpr int bq_waves_per_simd(int nb) { return nb <= 16 ? 4 : 2; }
// partition bounds in flight per thread
constexpr int bq_unroll(int nb) { return nb <= 8 ? 2 : 1; }
constexpr int BC_BS = 256;               // chunk-scoring workgroup

struct CutArgs {
    const BatchQuery *q;
    const BoundArgs *qb;
    TreeArgs t;
    uint32_t G, H, L;
    uint32_t d;
    const uint32_t *poff, *psz;
    const float *cstat;  // element e of chunk c at [e * nchunks + c]
    uint64_t nchunks;
    uint2 *tasks;        // unit u: tasks[u * ucap ...] = (partition, first chunk index within the unit)
    uint64_t ucap;
    uint2 *ucnt;         // unit u: (entries, chunks)
    unsigned int *ctr;
    const uint32_t *qflag;  // second pass: settled queries are skipped
    uint32_t S;             // slices per query (gridDim.y)
    const uint32_t *flist;  // second pass: the flagged queries ([0] count); workgroup x takes entries x, x + gridDim.x, ...
};

// Index i < n of the last entry with off[i] <= t (off non-decreasing, off[0] = 0).
__device__ __forceinline__ uint32_t last_le(const uint32_t *off, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Exclusive scan of cnt[0..n) (n <= BQ_PCAP) into off[0..n], BQ_EPT
// consecutive entries per thread.  All threads call it.
__device__ __forceinline__ void bq_scan(const uint32_t *cnt, uint32_t *off, uint32_t n, uint32_t *wsum,
                                        uint32_t *tot) {
    const uint32_t i0 = BQ_EPT * threadIdx.x;
    uint32_t c[BQ_EPT], mine = 0;
#pragma unroll
    for (uint32_t e = 0; e < BQ_EPT; e++) {
        c[e] = i0 + e < n ? cnt[i0 + e] : 0u;
        mine += c[e];
    }
    uint32_t run = block_excl_scan<BQ_BS>(mine, wsum, tot);
#pragma unroll
    for (uint32_t e = 0; e < BQ_EPT; e++) {
        if (i0 + e < n) off[i0 + e] = run;
        run += c[e];
    }
    if (threadIdx.x == 0) off[n] = *tot;
}

// Cuts of the np listed partitions lp[0..np) (positions lpos) and their
// task entries at out[nent ...], chunk indices from nch0 on; all threads
// call it.  Returns the chunks listed.
template <int NB>
__device__ uint32_t bq_cut_listed(const CutArgs &a, const BoundArgs &b, float tau, const uint32_t *lp,
                                  const uint32_t *lpos, uint32_t np, uint32_t *lcnt, uint32_t *loff, uint32_t *wsum,
                                  uint32_t *tot, uint2 *out, uint32_t nent, uint32_t nch0) {
    const uint32_t tid = threadIdx.x;
    bq_scan(lcnt, loff, np, wsum, tot);
    __syncthreads();
    // suffix-row bounds, R rows per thread in flight
    constexpr int R = bq_unroll(NB);
    const uint32_t nch = loff[np];
    const uint64_t S = a.nchunks;
    for (uint32_t t0 = tid; t0 < nch; t0 += R * BQ_BS) {
        float ub[R];
        uint32_t ii[R], cc[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t t = t0 + r * BQ_BS;
            ub[r] = INFINITY;
            ii[r] = 0;
            cc[r] = 0;
            if (t < nch) {
                ii[r] = last_le(loff, np, t);
                cc[r] = t - loff[ii[r]];
                const float *row = a.cstat + (lpos[ii[r]] / PTA_CHUNK + cc[r]);  // element e at row[e * S]
                ub[r] = lp_bound<NB>(b, row, row + a.d * S, row[2 * a.d * S], row[(2 * a.d + 1) * S], S);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (ub[r] < tau) atomicMin(&lcnt[ii[r]], cc[r]);
    }
    __syncthreads();
    bq_scan(lcnt, loff, np, wsum, tot);
    __syncthreads();
    for (uint32_t i = tid; i < np; i += BQ_BS) out[nent + i] = make_uint2(lp[i], nch0 + loff[i]);
    const uint32_t listed = loff[np];
    __syncthreads();  // the caller reuses the lists
    return listed;
}

template <int NB>
__global__ __launch_bounds__(BQ_BS, bq_waves_per_simd(NB)) void batch_cut_kernel(CutArgs a) {
    constexpr int U = bq_unroll(NB);
    __shared__ uint16_t glist[BMASK_PAIRS];
    __shared__ uint32_t lp[BQ_PCAP], lpos[BQ_PCAP], lcnt[BQ_PCAP], loff[BQ_PCAP + 1];
    __shared__ uint32_t wsum[BQ_BS / 64];
    __shared__ uint32_t ng, np, tot;
    const uint32_t sl = blockIdx.y, tid = threadIdx.x;
    const uint32_t nent_list = a.flist ? a.flist[0] : gridDim.x;
    for (uint32_t f = blockIdx.x; f < nent_list; f += gridDim.x) {
    const uint32_t q = a.flist ? a.flist[1 + f] : f;
    const uint64_t u = (uint64_t)f * a.S + sl;
    if (a.qflag && a.qflag[q] == QF_DONE) {
        if (tid == 0) a.ucnt[u] = make_uint2(0u, 0u);
        continue;
    }
    const BoundArgs &b = a.qb[q];
    const float tau = a.q[q].tau;
    const bool sat = at_smax(a.q[q]);
    const uint32_t tau_id = a.q[q].tau_id;
    if (tid == 0) {
        ng = 0;
        np = 0;
    }
    __syncthreads();
    // groups of this slice
    for (uint32_t g = sl + a.S * tid; g < a.G; g += a.S * BQ_BS) {
        const float ub = node_bound<NB>(b, a.t, a.G + g);
        if (ub > -INFINITY && ub >= tau && !(sat && a.t.tminid[a.G + g] > tau_id))
            glist[atomicAdd(&ng, 1u)] = (uint16_t)g;
    }
    __syncthreads();
    const uint32_t H = a.H, per = 1u << H, total = ng << H;
    constexpr uint32_t STEP = BQ_BS * U;  // partition bounds per step
    uint2 *out = a.tasks + u * a.ucap;
    uint32_t parts = 0, nent = 0, nchk = 0;
    auto flush = [&]() {
        const uint32_t n = np;
        for (uint32_t i = tid; i < n; i += BQ_BS) {
            const uint32_t p = lp[i];
            lpos[i] = a.poff[p];
            lcnt[i] = (a.psz[p] + PTA_CHUNK - 1) / PTA_CHUNK;
        }
        __syncthreads();
        nchk += bq_cut_listed<NB>(a, b, tau, lp, lpos, n, lcnt, loff, wsum, &tot, out, nent, nchk);
        nent += n;
        parts += n;
        if (tid == 0) np = 0;
        __syncthreads();
    };
    for (uint32_t i0 = 0; i0 < total; i0 += STEP) {
        if (np > BQ_PCAP - STEP) flush();  // uniform: np is read after a barrier
        uint32_t p[U];
        bool in[U];
#pragma unroll
        for (int v = 0; v < U; v++) {
            const uint32_t i = i0 + v * BQ_BS + tid;
            p[v] = i < total ? ((uint32_t)glist[i >> H] << H) + (i & (per - 1)) : a.L;
            const float ub = p[v] < a.L ? node_bound<NB>(b, a.t, a.t.p2 + p[v]) : -INFINITY;
            in[v] = ub > -INFINITY && ub >= tau && !(sat && a.t.tminid[a.t.p2 + p[v]] > tau_id);
        }
#pragma unroll
        for (int v = 0; v < U; v++)
            if (in[v]) lp[atomicAdd(&np, 1u)] = p[v];
        __syncthreads();
    }
    if (np) flush();
    if (tid == 0) {
        a.ucnt[u] = make_uint2(nent, nchk);
        if (parts) atomicAdd(reinterpret_cast<unsigned long long *>(a.ctr + 2), (unsigned long long)parts);
    }
    __syncthreads();  // the next entry reuses the lists
    }
}

// Exclusive prefix of the units' chunk counts into uoff[0..nu] (nu =
// nu_host, or the flagged-list count x S); one workgroup of 1024 threads,
// contiguous ranges per thread, 16 loads in flight.
__global__ __launch_bounds__(1024) void batch_unit_scan_kernel(const uint2 *__restrict__ ucnt, uint32_t nu_host,
                                                               const uint32_t *__restrict__ flist, uint32_t S,
                                                               uint32_t *__restrict__ uoff) {
    __shared__ uint32_t wsum[1024 / 64];
    __shared__ uint32_t tot;
    const uint32_t nu = flist ? flist[0] * S : nu_host;
    const uint32_t per = (nu + 1023) / 1024, b0 = threadIdx.x * per, b1 = b0 + per < nu ? b0 + per : nu;
    uint32_t mine = 0;
    for (uint32_t i0 = b0; i0 < b1; i0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = i0 + r < b1 ? ucnt[i0 + r].y : 0u;
#pragma unroll
        for (int r = 0; r < 16; r++) mine += v[r];
    }
    uint32_t run = block_excl_scan<1024>(mine, wsum, &tot);
    for (uint32_t i0 = b0; i0 < b1; i0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = i0 + r < b1 ? ucnt[i0 + r].y : 0u;
#pragma unroll
        for (int r = 0; r < 16; r++)
            if (i0 + r < b1) {
                uoff[i0 + r] = run;
                run += v[r];
            }
    }
    if (threadIdx.x == 0) uoff[nu] = tot;
}

struct ChunkArgs {
    const BatchQuery *q;
    const float *pcols;
    uint64_t ld;
    const uint32_t *poff, *psz, *pid;
    const uint2 *tasks;
    uint64_t ucap;
    const uint2 *ucnt;
    const uint32_t *uoff;
    uint32_t nu, S;            // units (nu_host) and slices per query
    const uint32_t *flist;     // second pass: unit u belongs to query flist[1 + u / S] ([0] count)
    uint64_t *cand;
    unsigned int *ccount;
    unsigned long long *scored;  // per workgroup
    const uint32_t *cminid;      // smallest object id per chunk
};

constexpr uint32_t BC_STAGE = 256;  // candidate keys a wave stages in LDS before one append

__global__ __launch_bounds__(BC_BS) void batch_chunk_kernel(ChunkArgs a) {
    __shared__ unsigned long long wpairs[BC_BS / 64][2];
    __shared__ uint64_t stage[BC_BS / 64][BC_STAGE];
    // the wave index made visibly uniform: the whole chunk cursor (unit,
    // entry, partition, query) then lives in scalar registers and its loads
    // are scalar (62 vector registers instead of 76: eight waves per SIMD)
    const uint32_t lane = lane_id(), wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const uint32_t nwaves = gridDim.x * (BC_BS / 64), gw = blockIdx.x * (BC_BS / 64) + wave;
    const uint32_t nu = a.flist ? a.flist[0] * a.S : a.nu;
    const uint32_t total = a.uoff[nu];
    const uint32_t per = (total + nwaves - 1) / nwaves;
    uint32_t t = gw * per;
    const uint32_t t_end = t + per < total ? t + per : total;
    uint32_t pairs = 0;
    unsigned long long terms = 0;
    uint64_t *wst = stage[wave];
    if (t < t_end) {
        const uint32_t *uoff = a.uoff;
        uint32_t u = wave_last_le([&](uint32_t i) { return uoff[i]; }, 0u, nu, t);
        uint32_t ub = uoff[u], ue = uoff[u + 1];
        const uint2 *tl = a.tasks + (uint64_t)u * a.ucap;
        uint32_t ne = a.ucnt[u].x;
        uint32_t j = wave_last_le([&](uint32_t i) { return tl[i].y; }, 0u, ne, t - ub);
        auto query_of = [&](uint32_t uu) { return a.flist ? a.flist[1 + uu / a.S] : uu / a.S; };
        uint32_t q = query_of(u);
        // the query's terms, lane j holding term j
        float lw = 0.0f, tau = 0.0f;
        uint32_t la = 0, tau_id = 0, m = 0;
        bool sat = false;
        auto load_query = [&]() {
            const BatchQuery &Q = a.q[q];
            lw = lane < TOPK_MAX_DIMS ? Q.w[lane] : 0.0f;
            la = lane < TOPK_MAX_DIMS ? Q.attr[lane] : 0u;
            tau = Q.tau;
            tau_id = Q.tau_id;
            m = Q.m;
            sat = at_smax(Q);
        };
        load_query();
        // staged candidates of query q (wave-uniform count); one append each flush
        uint32_t nst = 0;
        auto flush = [&]() {
            if (!nst) return;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&a.ccount[q], nst);
            base = __shfl(base, 0, 64);
            uint64_t *out = a.cand + (uint64_t)q * BATCH_CAP;
            for (uint32_t i = lane; i < nst; i += 64)
                if (base + i < BATCH_CAP) out[base + i] = wst[i];
            nst = 0;
            __builtin_amdgcn_wave_barrier();  // the stage is rewritten after every lane read it
        };
        uint32_t jb = tl[j].y, je = j + 1 < ne ? tl[j + 1].y : ue - ub;
        uint32_t p = tl[j].x;
        for (; t < t_end; t++) {
            while (t >= ue) {  // next unit with chunks
                u++;
                ub = ue;
                ue = uoff[u + 1];
                if (t < ue) {
                    tl = a.tasks + (uint64_t)u * a.ucap;
                    ne = a.ucnt[u].x;
                    j = 0;
                    jb = 0;
                    je = ne > 1 ? tl[1].y : ue - ub;
                    p = tl[0].x;
                    const uint32_t qn = query_of(u);
                    if (qn != q) {
                        flush();
                        q = qn;
                        load_query();
                    }
                }
            }
            const uint32_t lt = t - ub;
            while (lt >= je) {  // next entry with chunks
                j++;
                jb = je;
                je = j + 1 < ne ? tl[j + 1].y : ue - ub;
                p = tl[j].x;
            }
            const uint32_t c = lt - jb, sz = a.psz[p];
            // at the largest score, a chunk whose ids all exceed tau_id holds no candidate
            if (sat && a.cminid[a.poff[p] / PTA_CHUNK + c] > tau_id) continue;
            const uint32_t first = c * PTA_CHUNK + 4 * lane;
            const uint64_t pos0 = (uint64_t)a.poff[p] + first;
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t j0 = 0; j0 < m; j0 += 4) {
                f4 x[4];
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (j0 + v < m) {
                        const uint32_t at = (uint32_t)__builtin_amdgcn_readlane((int)la, (int)(j0 + v));
                        x[v] = *reinterpret_cast<const f4 *>(a.pcols + (uint64_t)at * a.ld + pos0);
                    }
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (j0 + v < m) {
                        const float w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lw), (int)(j0 + v)));
                        f4 pr = x[v] * w;
                        acc = acc + pr;
                    }
            }
            const uint32_t nvalid = first < sz ? (sz - first < 4u ? sz - first : 4u) : 0u;
            pairs += nvalid;
            terms += (unsigned long long)nvalid * m;
            uint64_t kk[4];
            uint32_t nin = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                bool in = (uint32_t)e < nvalid && acc[e] >= tau;
                uint32_t id = 0;
                if (in) {
                    id = a.pid[pos0 + e];
                    if (acc[e] == tau && id > tau_id) in = false;
                }
                kk[e] = in ? rkey(acc[e], id) : 0ull;
                nin += in ? 1u : 0u;
            }
            if (__ballot(nin != 0)) {
                const uint32_t incl = wave_incl_scan(nin);
                const uint32_t tot = __shfl(incl, 63, 64);
                if (nst + tot > BC_STAGE) flush();
                uint32_t pos = nst + incl - nin;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if (kk[e]) {
                        wst[pos] = kk[e];
                        pos++;
                    }
                }
                nst += tot;
            }
        }
        flush();
    }
    unsigned long long s = pairs, ts = terms;
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        ts += __shfl_xor(ts, o, 64);
    }
    if (lane == 0) {
        wpairs[wave][0] = s;
        wpairs[wave][1] = ts;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BC_BS / 64; w++) {
            s += wpairs[w][0];
            ts += wpairs[w][1];
        }
        a.scored[blockIdx.x] = s;
        a.scored[BATCH_MAX_GRID + blockIdx.x] = ts;
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
// Slices per query of the query-major pass: 1 in a first pass (every
// query its own workgroup), up to BQ_MAX_S in a second (few queries).
static uint32_t bq_slices(uint32_t nq, bool second, int first = 0) {
    if (second) return std::max(1u, std::min(BQ_MAX_S, BATCH_MAX_GRID / nq));
    return std::max(1u, std::min((uint32_t)(first ? first : 1), std::min(BQ_MAX_S, BATCH_MAX_GRID / nq)));
}

// Task-list entries both passes need (a unit lists at most its groups'
// partitions); 0 when they exceed the memory budget.
static uint64_t bq_task_entries(uint32_t nq, uint32_t p2, uint32_t H) {
    const uint32_t G = p2 >> H;
    uint64_t need = 0;
    for (int second = 0; second < 2; second++) {
        for (uint32_t S : {bq_slices(nq, second != 0), bq_slices(nq, second != 0, (int)BQ_MAX_S)})
            need = std::max(need, (uint64_t)nq * S * ((uint64_t)((G + S - 1) / S) << H));
    }
    constexpr uint64_t BUDGET = 4ull << 30;  // bytes
    return need * sizeof(uint2) <= BUDGET ? need : 0;
}

struct BatchVariant {
    int vpt, bs, qpi, digit, grid;
    int item;   // objects per work item (a multiple of PTA_CHUNK)
    int fused;  // 1: the query-major fused pass replaces mask, compaction and item scoring
};

using GraphKey = std::array<uint64_t, 17>;

// One batch in flight: its stream, pinned staging, device inputs and
// outputs, the scratch its kernels use, and the event that marks its
// results on the host.  The two slots share nothing a kernel writes, so
// their batches run concurrently on their own streams.
struct HostSlot {
    hipStream_t stream = nullptr;    // pipeline lane `slot`'s stream (owned by the context)
    unsigned char *h_in = nullptr, *h_out = nullptr;
    size_t in_cap = 0, out_cap = 0;
    unsigned char *d_in = nullptr;   // raw queries | taus | output descriptor | BatchQuery[nq] | BoundArgs[nq]
    unsigned char *d_out = nullptr;  // ids | scores | thr | status | cnt | flags | ctr
    size_t din_cap = 0, dout_cap = 0;
    hipEvent_t done = nullptr;       // recorded after the batch's results reached the host
    bool busy = false;
    bool direct = false;             // ids / scores / thresholds were written straight into the caller's arrays
    uint64_t ticket = 0;
    uint32_t nq = 0, ks = 0;
    uint64_t *out_ids = nullptr;
    float *out_scores = nullptr, *out_thr = nullptr;
    double t0 = 0;
    // scratch of this slot's kernels
    uint32_t nq_cap = 0;
    uint64_t plist_cap = 0, pcnt_cap = 0, item_cap = 0;
    uint16_t *d_plist = nullptr;     // per-partition query lists (nq entries each)
    unsigned int *d_pcnt = nullptr;  // per-partition list length
    BatchItem *d_items = nullptr;
    unsigned int *d_ctr = nullptr;
    unsigned long long *d_scored = nullptr;  // per scoring workgroup: pairs, then terms scored
    // query-major pass: task lists per (query, slice) unit, their counts and
    // the prefix of their chunk counts (null when over the memory budget)
    uint2 *d_tasks = nullptr;
    uint64_t tasks_cap = 0;
    uint2 *d_ucnt = nullptr;
    uint32_t *d_uoff = nullptr;
    uint32_t unit_cap = 0;
    unsigned int *d_flist = nullptr;  // queries the first pass flagged for the second: [0] count, then ids
    unsigned int *d_ccount = nullptr;
    uint64_t *d_cand = nullptr;
};

// True when p is pinned (page-locked) host memory the GPU can write.
static bool pinned_host_query(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t at;
    const hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // pageable memory: an expected miss, not an error to report later
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

// Waits for an event by polling (a blocking wait can add tens of
// microseconds of wake-up latency per batch), then confirms it.
static hipError_t wait_event(hipEvent_t e) {
    for (int i = 0; i < (1 << 22); i++) {
        const hipError_t q = hipEventQuery(e);
        if (q != hipErrorNotReady) return q;
    }
    return hipEventSynchronize(e);
}

struct BatchState {
    uint64_t gen = ~0ull;   // layout generation the buffers and graphs belong to
    // Two batches may be in flight (topk_batch_submit), each in its own slot.
    HostSlot slot[2];
    int cur = 0;                     // slot of the batch being enqueued
    int next = 0;                    // slot the next submit uses
    uint64_t next_ticket = 1;
    int cus = 0;
    int chunk_wgs = 0;               // batch_chunk_kernel workgroups resident per CU (occupancy query)
    std::map<GraphKey, hipGraphExec_t> graphs;
    // pinned-memory answers for recently seen output arrays (callers reuse them)
    std::map<const void *, bool> pinned;
};

static bool pinned_host(BatchState &B, const void *p) {
    auto it = B.pinned.find(p);
    if (it != B.pinned.end()) return it->second;
    if (B.pinned.size() > 64) B.pinned.clear();
    const bool r = pinned_host_query(p);
    B.pinned[p] = r;
    return r;
}

// Staged inputs: the caller's topk_query records and given thresholds are
// copied from pinned memory (the first `host` bytes); the device builds the
// BatchQuery / BoundArgs forms behind them.
struct InLayout {
    size_t raw_off, tau_off, od_off, host, q_off, b_off, bytes;
    static InLayout make(uint32_t nq) {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        InLayout l;
        l.raw_off = 0;
        l.tau_off = al(sizeof(topk_query) * nq);
        l.od_off = al(l.tau_off + sizeof(float) * nq);
        l.host = l.od_off + sizeof(OutDesc);
        l.q_off = al(l.host);
        l.b_off = al(l.q_off + sizeof(BatchQuery) * nq);
        l.bytes = l.b_off + sizeof(BoundArgs) * nq;
        return l;
    }
};

struct OutLayout {
    size_t ids_off, sc_off, thr_off, st_off, cnt_off, fl_off, ctr_off, ctr2_off, bytes;
    static OutLayout make(uint32_t nq, uint32_t ks) {
        OutLayout l;
        l.ids_off = 0;
        l.sc_off = l.ids_off + sizeof(uint64_t) * nq * ks;
        l.thr_off = l.sc_off + ((sizeof(float) * nq * ks + 15) & ~(size_t)15);
        l.st_off = l.thr_off + sizeof(float) * nq;
        l.cnt_off = l.st_off + sizeof(uint32_t) * nq;
        l.fl_off = l.cnt_off + sizeof(uint32_t) * nq;
        l.ctr_off = (l.fl_off + sizeof(uint32_t) * nq + 15) & ~(size_t)15;
        l.ctr2_off = l.ctr_off + sizeof(uint32_t) * BCTR_WORDS;
        l.bytes = l.ctr2_off + sizeof(uint32_t) * BCTR_WORDS;
        return l;
    }
};

static void drop_graphs(BatchState &B) {
    for (auto &g : B.graphs) hipGraphExecDestroy(g.second);
    B.graphs.clear();
}

void batch_free(topk_ctx *ctx) {
    BatchState *B = static_cast<BatchState *>(ctx->batch);
    if (!B) return;
    for (HostSlot &h : B->slot)  // batches still in flight
        if (h.stream) hipStreamSynchronize(h.stream);
    drop_graphs(*B);
    for (HostSlot &h : B->slot) {
        if (h.h_in) hipHostFree(h.h_in);
        if (h.h_out) hipHostFree(h.h_out);
        hipFree(h.d_in);
        hipFree(h.d_out);
        hipFree(h.d_plist);
        hipFree(h.d_tasks);
        hipFree(h.d_ucnt);
        hipFree(h.d_uoff);
        hipFree(h.d_flist);
        hipFree(h.d_pcnt);
        hipFree(h.d_items);
        hipFree(h.d_ctr);
        hipFree(h.d_scored);
        hipFree(h.d_ccount);
        hipFree(h.d_cand);
        if (h.done) hipEventDestroy(h.done);
    }
    delete B;
    ctx->batch = nullptr;
}

template <class T>
static int grow_dev(T *&p, uint64_t &cap, uint64_t need, bool &changed) {
    if (need <= cap) return TOPK_OK;
    hipFree(p);
    p = nullptr;
    cap = 0;
    changed = true;
    if (hipMalloc(&p, sizeof(T) * need) != hipSuccess) return TOPK_ERR_NOMEM;
    cap = need;
    return TOPK_OK;
}

// Height of the mask's groups: 64 partitions, fewer on small layouts.
static uint32_t group_height(uint32_t p2) {
    uint32_t h = 0;
    while ((1u << (h + 1)) <= BMASK_MAX_PER && (1u << (h + 1)) <= p2) h++;
    return h;
}

static int batch_drain(topk_ctx *ctx, BatchState &B);

// Sizes the shared device buffers and slot si's staging for nq queries.
// Growing anything first completes the batches in flight (they use the
// buffers and the captured graphs).
static int ensure(topk_ctx *ctx, BatchState &B, uint32_t nq, uint32_t ks, int si) {
    const PtaLayout &P = ctx->pta;
    HostSlot &H = B.slot[si];
    const uint32_t L = P.nparts, nw = (nq + 63) / 64;
    const size_t in_need = InLayout::make(nq).bytes, out_need = OutLayout::make(nq, ks).bytes;
    const uint64_t items_need =
        (uint64_t)L * (4 * nw + 1) * ((P.maxsz + ITEM_MIN_OBJECTS - 1) / ITEM_MIN_OBJECTS + 1);
    const uint64_t te = bq_task_entries(nq, P.p2, group_height(P.p2));
    const bool grow = B.gen != P.generation || nq > H.nq_cap || in_need > H.din_cap || out_need > H.dout_cap ||
                      in_need > H.in_cap || out_need > H.out_cap || (uint64_t)L * nq > H.plist_cap ||
                      (uint64_t)L > H.pcnt_cap || items_need > H.item_cap || !H.d_ctr || !H.done ||
                      te > H.tasks_cap || (te && nq * BQ_MAX_S > H.unit_cap) || (!te && H.d_tasks);
    if (!grow) return TOPK_OK;
    int rc = batch_drain(ctx, B);
    if (rc) return rc;
    bool changed = false;
    if (B.gen != P.generation) {
        changed = true;
        B.gen = P.generation;
    }
    if (nq > H.nq_cap) {
        uint32_t cap = 64;
        while (cap < nq) cap *= 2;
        hipFree(H.d_ccount);
        hipFree(H.d_cand);
        hipFree(H.d_flist);
        H.d_ccount = nullptr;
        H.d_cand = nullptr;
        H.d_flist = nullptr;
        H.nq_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&H.d_flist, sizeof(unsigned int) * (cap + 1)));
        TK_HIP(hipMalloc(&H.d_ccount, sizeof(unsigned int) * cap));
        TK_HIP(hipMalloc(&H.d_cand, sizeof(uint64_t) * BATCH_CAP * cap));
        H.nq_cap = cap;
    }
    if (in_need > H.din_cap) {
        hipFree(H.d_in);
        H.d_in = nullptr;
        H.din_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&H.d_in, in_need));
        H.din_cap = in_need;
    }
    if (out_need > H.dout_cap) {
        hipFree(H.d_out);
        H.d_out = nullptr;
        H.dout_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&H.d_out, out_need));
        H.dout_cap = out_need;
    }
    if (in_need > H.in_cap) {
        if (H.h_in) hipHostFree(H.h_in);
        H.h_in = nullptr;
        H.in_cap = 0;
        changed = true;
        TK_HIP(hipHostMalloc(&H.h_in, in_need, hipHostMallocDefault));
        H.in_cap = in_need;
    }
    if (out_need > H.out_cap) {
        if (H.h_out) hipHostFree(H.h_out);
        H.h_out = nullptr;
        H.out_cap = 0;
        changed = true;
        TK_HIP(hipHostMalloc(&H.h_out, out_need, hipHostMallocDefault));
        H.out_cap = out_need;
    }
    if (!H.done) TK_HIP(hipEventCreateWithFlags(&H.done, hipEventDisableTiming));
    // slot i runs on pipeline lane i's stream: no streams beyond the
    // context's, so the two slots keep hardware queues of their own
    H.stream = ctx->lane[si].stream;
    rc = grow_dev(H.d_plist, H.plist_cap, (uint64_t)L * nq, changed);
    if (!rc) rc = grow_dev(H.d_pcnt, H.pcnt_cap, (uint64_t)L, changed);
    // budget >= 16 queries, chunks >= ITEM_MIN_OBJECTS objects
    if (!rc) rc = grow_dev(H.d_items, H.item_cap, items_need, changed);
    if (rc) return rc;
    {
        const uint32_t units = nq * BQ_MAX_S;
        if (!te && H.d_tasks) {  // over the budget for this batch size: the item path runs
            hipFree(H.d_tasks);
            H.d_tasks = nullptr;
            H.tasks_cap = 0;
            changed = true;
        }
        if (te) {
            rc = grow_dev(H.d_tasks, H.tasks_cap, te, changed);
            if (!rc && units > H.unit_cap) {
                hipFree(H.d_ucnt);
                hipFree(H.d_uoff);
                H.d_ucnt = nullptr;
                H.d_uoff = nullptr;
                H.unit_cap = 0;
                changed = true;
                TK_HIP(hipMalloc(&H.d_ucnt, sizeof(uint2) * units));
                TK_HIP(hipMalloc(&H.d_uoff, sizeof(uint32_t) * (units + 1)));
                H.unit_cap = units;
            }
            if (rc) return rc;
        }
    }
    if (!H.d_ctr) {
        changed = true;
        TK_HIP(hipMalloc(&H.d_ctr, sizeof(unsigned int) * BCTR_WORDS));
        TK_HIP(hipMalloc(&H.d_scored, sizeof(unsigned long long) * 2 * BATCH_MAX_GRID));
    }
    if (!B.cus) {
        TK_HIP(hipDeviceGetAttribute(&B.cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (B.cus < 1) B.cus = 1;
        // the chunk kernel splits its chunks evenly over a grid that must be
        // resident at once: as many workgroups per CU as its registers
        // allow, at most six, so two wave slots per SIMD stay free for the
        // other batch in flight (six measured 3% above eight)
        int nb = 0;
        TK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_chunk_kernel, BC_BS, 0));
        B.chunk_wgs = nb < 1 ? 1 : (nb > 6 ? 6 : nb);
    }
    if (changed) drop_graphs(B);
    return TOPK_OK;
}

static int dims_class(uint32_t d) { return d <= 8 ? 8 : (d <= 16 ? 16 : 32); }

// Objects of one work item: the variant's item size, at least one scoring
// pass of the shape and ITEM_MIN_OBJECTS (which bounds the item list).
static uint32_t item_objects(const BatchVariant &v) {
    uint32_t o = (uint32_t)(v.bs * v.vpt);
    o = o > (uint32_t)v.item ? o : (uint32_t)v.item;
    return o > ITEM_MIN_OBJECTS ? o : ITEM_MIN_OBJECTS;
}

static BatchVariant pick_variant(const topk_ctx *ctx, const BatchState &B) {
    const int D = dims_class(ctx->d);
    BatchVariant v;
    v.vpt = ctx->tune.batch_vpt ? ctx->tune.batch_vpt : (D == 8 ? 8 : (D == 16 ? 4 : 2));
    v.bs = ctx->tune.batch_bs ? ctx->tune.batch_bs : 256;
    v.qpi = ctx->tune.batch_qpi ? ctx->tune.batch_qpi : 128;
    v.digit = ctx->tune.batch_digit ? ctx->tune.batch_digit : 11;
    v.grid = ctx->tune.batch_grid ? ctx->tune.batch_grid : B.cus * (v.bs == 256 ? 4 : 2);
    v.item = ctx->tune.batch_item ? ctx->tune.batch_item : 4096;
    v.fused = ctx->tune.batch_fused == 2 ? 1 : 0;
    return v;
}

static bool variant_ok(int D, int vpt, int bs) {
    // 1024-thread workgroups (16 waves, one CU) need <= 128 VGPRs a thread
    if (bs == 1024) return D == 8 && vpt == 4;
    if (bs != 256 && bs != 512) return false;
    if (D == 8) return vpt == 4 || vpt == 8 || vpt == 16;
    if (D == 16) return vpt == 2 || vpt == 4 || vpt == 8;
    return vpt == 2 || vpt == 4;
}

static void launch_score(int D, const BatchVariant &v, hipStream_t s, const ScoreArgs &a) {
#define TK_SC(DD, VV, BB)                                                                                      \
    if (D == DD && v.vpt == VV && v.bs == BB) {                                                                \
        hipLaunchKernelGGL((batch_score_kernel<DD, VV, BB>), dim3(v.grid), dim3(BB), 0, s, a);                 \
        return;                                                                                                \
    }
    TK_SC(8, 4, 256) TK_SC(8, 8, 256) TK_SC(8, 16, 256) TK_SC(8, 4, 512) TK_SC(8, 8, 512) TK_SC(8, 16, 512)
    TK_SC(8, 4, 1024)
    TK_SC(16, 2, 256) TK_SC(16, 4, 256) TK_SC(16, 8, 256) TK_SC(16, 2, 512) TK_SC(16, 4, 512) TK_SC(16, 8, 512)
    TK_SC(32, 2, 256) TK_SC(32, 4, 256) TK_SC(32, 2, 512) TK_SC(32, 4, 512)
#undef TK_SC
}

// Enqueues one scoring pass: threshold stage (first pass; a second pass
// only for seeded reruns), two-level mask, compaction, scoring, final
// select.  The second pass works only on the queries the first one flagged.
static int enqueue_pass(topk_ctx *ctx, BatchState &B, hipStream_t s, uint32_t nq, uint32_t ks, int mode,
                        const BatchVariant &v, bool second) {
    const PtaLayout &P = ctx->pta;
    const uint32_t L = P.nparts;
    const InLayout il = InLayout::make(nq);
    const OutLayout ol = OutLayout::make(nq, ks);
    unsigned char *d_in = B.slot[B.cur].d_in, *d_out = B.slot[B.cur].d_out;
    BatchQuery *dq = reinterpret_cast<BatchQuery *>(d_in + il.q_off);
    BoundArgs *db = reinterpret_cast<BoundArgs *>(d_in + il.b_off);
    uint32_t *qflag = reinterpret_cast<uint32_t *>(d_out + ol.fl_off);
    const uint32_t *pass_flags = second ? qflag : nullptr;
    // counters: cleared by the prep kernel for the first pass, by the first
    // pass's final kernel for the second
    const TreeArgs ta{P.tmax, P.tmin, P.tsum, P.tminsum, P.tminid, P.p2, ctx->d};
    // a second-pass seed only serves learned misses that rerun with seeds
    const bool seed2 = mode == BMODE_LEARNED && ctx->tune.batch_retry == 2;
    hipEvent_t evt = nullptr;
    timing_begin(ctx, s, evt);
    if (second ? seed2 : mode == BMODE_SEED) {
        const int NB = dims_class(ctx->d);
        auto seed = NB == 8 ? batch_seed_kernel<8> : (NB == 16 ? batch_seed_kernel<16> : batch_seed_kernel<32>);
        hipLaunchKernelGGL(seed, dim3(nq), dim3(BSEED_BS), 0, s, dq, db, ta, L, P.pcols, P.ld, P.poff, P.psz, P.pid,
                           pass_flags, 0);
    } else if (mode == BMODE_LEARNED && !second) {
        const int rc = learned_tau_enqueue(ctx, s, dq, nq);
        if (rc) return rc;
        if (learned_saturated(ctx)) {
            // queries predicted at their largest score: a key from the
            // extreme-value bitmaps, else a seed (key) threshold
            if (P.xbits) {
                auto tk = ctx->d <= 16 ? batch_tiebits_kernel<16, 4> : batch_tiebits_kernel<32, 2>;
                hipLaunchKernelGGL(tk, dim3(nq), dim3(BXB_BS), 0, s, dq, P.xbits, ctx->d, P.nw, ctx->n, BXB_BUDGET);
            } else {  // no bitmaps (the layout was rebuilt after training): seeds
                const int NB = dims_class(ctx->d);
                auto seed = NB == 8 ? batch_seed_kernel<8> : (NB == 16 ? batch_seed_kernel<16> : batch_seed_kernel<32>);
                hipLaunchKernelGGL(seed, dim3(nq), dim3(BSEED_BS), 0, s, dq, db, ta, L, P.pcols, P.ld, P.poff, P.psz,
                                   P.pid, (const uint32_t *)nullptr, 1);
            }
        }
    }
    timing_end(ctx, s,
               second ? "batch_tau_second" : (mode == BMODE_SEED ? "batch_tau_seed" : (mode == BMODE_LEARNED ? "batch_tau_learned" : "batch_tau_given")),
               evt);
    const uint32_t H = group_height(P.p2), G = P.p2 >> H;
    // a query's group list and a group's partition counters live in LDS
    if (G > BMASK_PAIRS) return TOPK_ERR_UNSUPPORTED;  // layouts stop at PTA_MAX_PARTS partitions
    const int NB = dims_class(ctx->d);
    FinalArgs fa{dq,
                 B.slot[B.cur].d_cand,
                 B.slot[B.cur].d_ccount,
                 B.slot[B.cur].d_ctr,
                 ks,
                 reinterpret_cast<uint64_t *>(d_out + ol.ids_off),
                 reinterpret_cast<float *>(d_out + ol.sc_off),
                 reinterpret_cast<float *>(d_out + ol.thr_off),
                 reinterpret_cast<uint32_t *>(d_out + ol.st_off),
                 reinterpret_cast<uint32_t *>(d_out + ol.cnt_off),
                 reinterpret_cast<uint32_t *>(d_out + (second ? ol.ctr2_off : ol.ctr_off)),
                 qflag,
                 second ? 1 : 0,
                 B.slot[B.cur].d_scored,
                 0u,
                 seed2 ? 1 : 0,
                 B.slot[B.cur].d_pcnt,
                 L,
                 second ? nullptr : B.slot[B.cur].d_flist};
    if (v.fused && B.slot[B.cur].d_tasks) {
        // a second pass re-runs few queries: each is spread over S slices
        const uint32_t S = bq_slices(nq, second, ctx->tune.batch_slices);
        const uint64_t ucap = (uint64_t)((G + S - 1) / S) << H;
        // the second pass works through the list of flagged queries the
        // first final select wrote: a few workgroups per slice, not one per query
        const uint32_t *fl = second ? B.slot[B.cur].d_flist : nullptr;
        const uint32_t gx = second ? std::min(nq, 16u) : nq;
        const CutArgs ca{dq,       db,    ta,     G,       H,       L,         ctx->d,  P.poff, P.psz,
                         P.cstat,  (uint64_t)P.nchunks, B.slot[B.cur].d_tasks, ucap, B.slot[B.cur].d_ucnt, B.slot[B.cur].d_ctr, pass_flags, S, fl};
        auto ck = NB == 8 ? batch_cut_kernel<8> : (NB == 16 ? batch_cut_kernel<16> : batch_cut_kernel<32>);
        hipLaunchKernelGGL(ck, dim3(gx, S), dim3(BQ_BS), 0, s, ca);
        hipLaunchKernelGGL(batch_unit_scan_kernel, dim3(1), dim3(1024), 0, s, B.slot[B.cur].d_ucnt, nq * S, fl, S, B.slot[B.cur].d_uoff);
        const uint32_t g2 = (uint32_t)B.cus * (second ? 2 : (uint32_t)B.chunk_wgs);  // all resident at once
        const ChunkArgs xa{dq, P.pcols, P.ld, P.poff, P.psz, P.pid, B.slot[B.cur].d_tasks, ucap, B.slot[B.cur].d_ucnt, B.slot[B.cur].d_uoff, nq * S, S,
                           fl, B.slot[B.cur].d_cand, B.slot[B.cur].d_ccount, B.slot[B.cur].d_scored, P.cminid};
        hipEvent_t ev0 = nullptr;
        timing_begin(ctx, s, ev0);
        hipLaunchKernelGGL(batch_chunk_kernel, dim3(g2), dim3(BC_BS), 0, s, xa);
        timing_end(ctx, s, second ? "batch_score_second" : "batch_score", ev0);
        fa.nscored = g2;
        if (v.digit == 8)
            hipLaunchKernelGGL(batch_final_kernel<8>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
        else
            hipLaunchKernelGGL(batch_final_kernel<11>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
        TK_HIP(hipGetLastError());
        return TOPK_OK;
    }
    // mask layout: query-major (one workgroup per query) spreads a full
    // batch over the CUs; group-major (one per group and query slice) keeps
    // the few queries of a second pass from running one long chain each
    const int layout = ctx->tune.batch_mask ? ctx->tune.batch_mask : (second ? 2 : 1);
    if (layout == 1) {
        // slices per query: enough workgroups to cover the CUs on small
        // batches, and at most BMASK_PAIRS groups per workgroup
        uint32_t S = std::min(16u, std::max(1u, (4u * (uint32_t)B.cus + nq - 1) / nq));
        S = std::max(S, (G + BMASK_PAIRS - 1) / BMASK_PAIRS);
        auto mask = NB == 8 ? batch_mask_query_kernel<8>
                            : (NB == 16 ? batch_mask_query_kernel<16> : batch_mask_query_kernel<32>);
        hipLaunchKernelGGL(mask, dim3(nq, S), dim3(BMASK_BS), 0, s, dq, db, ta, G, H, L, nq, B.slot[B.cur].d_pcnt, B.slot[B.cur].d_plist,
                           pass_flags);
    } else {
        const uint32_t QS = BMASK_PAIRS >> H, Y = (nq + QS - 1) / QS;
        auto mask = NB == 8 ? batch_mask_kernel<8> : (NB == 16 ? batch_mask_kernel<16> : batch_mask_kernel<32>);
        hipLaunchKernelGGL(mask, dim3(G, Y), dim3(BMASK_BS), 0, s, dq, db, ta, G, H, L, nq, QS, B.slot[B.cur].d_pcnt,
                           B.slot[B.cur].d_plist, pass_flags);
    }
    const uint32_t ochunk = item_objects(v);
    hipLaunchKernelGGL(batch_compact_kernel, dim3((L + BAUX_BS - 1) / BAUX_BS), dim3(BAUX_BS), 0, s, B.slot[B.cur].d_pcnt, L,
                       (uint32_t)v.qpi, ochunk, B.slot[B.cur].item_cap, P.psz, B.slot[B.cur].d_items, B.slot[B.cur].d_ctr);
    ScoreArgs sa{P.pcols, P.ld, ctx->d,    P.poff,     P.psz,  P.pid,  B.slot[B.cur].d_items, B.slot[B.cur].item_cap,
                 ochunk,  B.slot[B.cur].d_plist, nq, B.slot[B.cur].d_ctr, dq, B.slot[B.cur].d_cand, B.slot[B.cur].d_ccount, P.cstat,  P.crow, (uint64_t)P.nchunks,
                 B.slot[B.cur].d_scored, P.cminid};
    BatchVariant vs = v;
    if (second) vs.grid = v.grid / 8 > 32 ? v.grid / 8 : 32;  // the second pass has few items
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    launch_score(dims_class(ctx->d), vs, s, sa);
    timing_end(ctx, s, second ? "batch_score_second" : "batch_score", ev0);
    fa.nscored = (uint32_t)vs.grid;
    if (v.digit == 8)
        hipLaunchKernelGGL(batch_final_kernel<8>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
    else
        hipLaunchKernelGGL(batch_final_kernel<11>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

// Enqueues the kernels of a whole batch on s (its inputs are on the
// device): prep, the first pass, the second pass for flagged queries.
static int batch_enqueue(topk_ctx *ctx, BatchState &B, hipStream_t s, uint32_t nq, uint32_t ks, int mode,
                         const BatchVariant &v) {
    const InLayout il = InLayout::make(nq);
    const OutLayout ol = OutLayout::make(nq, ks);
    HostSlot &H = B.slot[B.cur];
    unsigned char *d_in = H.d_in;
    BatchQuery *dq = reinterpret_cast<BatchQuery *>(d_in + il.q_off);
    BoundArgs *db = reinterpret_cast<BoundArgs *>(d_in + il.b_off);
    const uint32_t L = ctx->pta.nparts;
    uint32_t pg = (L + 63) / 64;
    if (pg > 1024) pg = 1024;
    if (pg < nq) pg = nq;
    // the staged queries, taus and output descriptor (pinned) -> device
    TK_HIP(hipMemcpyAsync(H.d_in, H.h_in, il.host, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(batch_prep_kernel, dim3(pg), dim3(64), 0, s,
                       reinterpret_cast<const topk_query *>(d_in + il.raw_off),
                       mode == BMODE_GIVEN ? reinterpret_cast<const float *>(d_in + il.tau_off) : nullptr, nq, ctx->n,
                       ctx->d, ctx->pta.tmax + 1, ctx->pta.tmin + 1, 2 * (uint64_t)ctx->pta.p2, dq, db, H.d_pcnt, L,
                       H.d_ccount, H.d_ctr, H.d_flist);
    int rc = enqueue_pass(ctx, B, s, nq, ks, mode, v, false);
    if (!rc) rc = enqueue_pass(ctx, B, s, nq, ks, mode, v, true);
    if (rc) return rc;
    const PublishArgs pa{H.d_out,  ol.ids_off, ol.sc_off, ol.thr_off, ol.st_off,
                         ol.bytes, (uint64_t)nq * ks, nq, reinterpret_cast<const OutDesc *>(d_in + il.od_off), H.h_out};
    hipLaunchKernelGGL(batch_publish_kernel, dim3(BPUB_GRID), dim3(256), 0, s, pa);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

static int batch_launch(topk_ctx *ctx, BatchState &B, hipStream_t s, uint32_t nq, uint32_t ks, int mode,
                        const BatchVariant &v) {
    if (!ctx->tune.batch_graph || ctx->timing_on) return batch_enqueue(ctx, B, s, nq, ks, mode, v);
    const GraphKey key{nq, ks, (uint64_t)mode, (uint64_t)v.vpt, (uint64_t)v.bs, (uint64_t)v.qpi, (uint64_t)v.digit,
                       (uint64_t)v.grid, ctx->d, mode == BMODE_LEARNED ? learned_version(ctx) : 0,
                       (uint64_t)ctx->tune.batch_mask, (uint64_t)B.cur, (uint64_t)v.item,
                       (uint64_t)ctx->tune.batch_retry, (uint64_t)ctx->tune.batch_retry_width, (uint64_t)v.fused,
                       (uint64_t)ctx->tune.batch_slices};
    auto it = B.graphs.find(key);
    if (it == B.graphs.end()) {
        TK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue(ctx, B, s, nq, ks, mode, v);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        TK_HIP(e);
        hipGraphExec_t ge = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        TK_HIP(ei);
        it = B.graphs.emplace(key, ge).first;
    }
    TK_HIP(hipGraphLaunch(it->second, s));
    return TOPK_OK;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace topk

int topk::make_batch_query(const topk_ctx *ctx, const topk_query &Q, BatchQuery &b) {
    if (Q.k < 1 || Q.qq < 1 || Q.qq > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    if (Q.k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    memset(&b, 0, sizeof(b));
    b.k = (uint32_t)(ctx->n < Q.k ? ctx->n : Q.k);
    b.kuser = Q.k;
    b.m = Q.qq;
    b.ascending = 1;
    for (uint32_t j = 0; j < Q.qq; j++) {
        if (Q.attr[j] >= ctx->d) return TOPK_ERR_INVALID;
        b.attr[j] = Q.attr[j];
        b.w[j] = Q.weights[Q.attr[j]];
        if (j && Q.attr[j] <= Q.attr[j - 1]) b.ascending = 0;
    }
    if (b.ascending)
        for (uint32_t j = 0; j < Q.qq; j++) b.wa[Q.attr[j]] = b.w[j];
    b.tau_id = 0xFFFFFFFFu;
    b.smax = INFINITY;  // not known here: never taken as the largest score
    const BoundArgs g = make_bound_args(ctx->d, Q.qq, Q.weights, Q.attr);
    for (uint32_t i = 0; i < g.nd && i < 16; i++) {
        b.ws[i] = (float)g.w[i];
        b.as[i] = g.attr[i];
    }
    return TOPK_OK;
}

namespace topk {

// Validates the queries and copies them into the pinned staging area.
static int stage_queries(topk_ctx *ctx, BatchState &B, uint32_t nq, const topk_query *qs, uint32_t ks) {
    for (uint32_t i = 0; i < nq; i++) {
        const topk_query &Q = qs[i];
        if (Q.k < 1 || Q.k > ks || Q.qq < 1 || Q.qq > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
        if (Q.k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
        for (uint32_t j = 0; j < Q.qq; j++)
            if (Q.attr[j] >= ctx->d) return TOPK_ERR_INVALID;
    }
    memcpy(B.slot[B.cur].h_in + InLayout::make(nq).raw_off, qs, sizeof(topk_query) * nq);
    return TOPK_OK;
}

static int resolve_mode(const topk_ctx *ctx) {
    if (ctx->tune.batch_mode == BMODE_SEED) return BMODE_SEED;
    if (ctx->tune.batch_mode == BMODE_LEARNED) return learned_ready(ctx) ? BMODE_LEARNED : -1;
    return learned_ready(ctx) ? BMODE_LEARNED : BMODE_SEED;
}

// Completes the batch of slot si: waits for its result copy, copies the
// results out, answers any query the second pass could not settle either
// through topk_fte, and sets the statistics.
static int batch_complete(topk_ctx *ctx, BatchState &B, int si) {
    HostSlot &H = B.slot[si];
    if (!H.busy) return TOPK_OK;
    H.busy = false;
    TK_HIP(wait_event(H.done));
    const uint32_t nq = H.nq, ks = H.ks;
    const topk_query *qs = reinterpret_cast<const topk_query *>(H.h_in + InLayout::make(nq).raw_off);
    const OutLayout ol = OutLayout::make(nq, ks);
    const uint64_t *ids = reinterpret_cast<const uint64_t *>(H.h_out + ol.ids_off);
    const float *sc = reinterpret_cast<const float *>(H.h_out + ol.sc_off);
    const float *thr = reinterpret_cast<const float *>(H.h_out + ol.thr_off);
    const uint32_t *st = reinterpret_cast<const uint32_t *>(H.h_out + ol.st_off);
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(H.h_out + ol.cnt_off);
    const uint32_t *fl = reinterpret_cast<const uint32_t *>(H.h_out + ol.fl_off);
    const uint32_t *c1 = reinterpret_cast<const uint32_t *>(H.h_out + ol.ctr_off);
    const uint32_t *c2 = reinterpret_cast<const uint32_t *>(H.h_out + ol.ctr2_off);
    uint64_t candidates = 0;
    uint32_t missed = 0, retried = 0;
    std::vector<uint32_t> fte;
    for (uint32_t i = 0; i < nq; i++) {
        candidates += cnt[i] < BATCH_CAP ? cnt[i] : BATCH_CAP;
        // a query counts once as a fallback: its first threshold missed, or
        // the device passes could not settle it
        missed += fl[i] == QF_SEED || fl[i] == QF_LOWER || st[i] ? 1u : 0u;
        retried += fl[i] == QF_GIVEN ? 1u : 0u;
        if (st[i]) fte.push_back(i);
    }
    // results are laid out [q * ks + j] on both sides; query i owns its
    // first k_i entries (min(k_i, n) results, then padding)
    bool full = ctx->n >= ks;
    for (uint32_t i = 0; full && i < nq; i++) full = qs[i].k == ks;
    if (H.direct) {
        // the result copy wrote the caller's arrays
    } else if (full) {  // every query fills its stride: one copy per array
        if (H.out_ids) memcpy(H.out_ids, ids, sizeof(uint64_t) * nq * ks);
        if (H.out_scores) memcpy(H.out_scores, sc, sizeof(float) * nq * ks);
    } else {
        for (uint32_t i = 0; i < nq; i++) {
            const uint32_t kq = qs[i].k, c = (uint32_t)(ctx->n < kq ? ctx->n : kq);
            const uint64_t o = (uint64_t)i * ks;
            if (H.out_ids) {
                memcpy(H.out_ids + o, ids + o, sizeof(uint64_t) * c);
                for (uint32_t j = c; j < kq; j++) H.out_ids[o + j] = ~0ull;
            }
            if (H.out_scores) {
                memcpy(H.out_scores + o, sc + o, sizeof(float) * c);
                for (uint32_t j = c; j < kq; j++) H.out_scores[o + j] = -INFINITY;
            }
        }
    }
    if (H.out_thr && !H.direct) memcpy(H.out_thr, thr, sizeof(float) * nq);
    for (uint32_t i : fte) {
        const topk_query &Q = qs[i];
        float t = 0;
        const int rc = topk_fte(ctx, Q.k, (uint8_t)Q.qq, Q.weights, Q.attr,
                                H.out_ids ? H.out_ids + (uint64_t)i * ks : nullptr,
                                H.out_scores ? H.out_scores + (uint64_t)i * ks : nullptr, &t);
        if (rc) return rc;
        if (H.out_thr) H.out_thr[i] = t;
    }
    topk_stats &S = ctx->stats;
    S.queries = nq;
    S.fallback = missed;
    S.retried = retried;
    S.fte_fallback = (uint32_t)fte.size();
    S.items = c1[0] + c1[1];
    S.candidates = candidates;
    S.partitions_scored = ((uint64_t)c1[2] | ((uint64_t)c1[3] << 32)) + ((uint64_t)c2[2] | ((uint64_t)c2[3] << 32));
    S.partitions_total = (uint64_t)ctx->pta.nonempty * nq;
    S.tuple_count = ((uint64_t)c1[4] | ((uint64_t)c1[5] << 32)) + ((uint64_t)c2[4] | ((uint64_t)c2[5] << 32));
    S.terms = ((uint64_t)c1[8] | ((uint64_t)c1[9] << 32)) + ((uint64_t)c2[8] | ((uint64_t)c2[9] << 32));
    S.ms_last = now_ms() - H.t0;
    return TOPK_OK;
}

// Completes the batches in flight with tickets <= upto, oldest first.
static int batch_wait_upto(topk_ctx *ctx, BatchState &B, uint64_t upto) {
    for (;;) {
        int si = -1;
        for (int i = 0; i < 2; i++)
            if (B.slot[i].busy && B.slot[i].ticket <= upto && (si < 0 || B.slot[i].ticket < B.slot[si].ticket)) si = i;
        if (si < 0) return TOPK_OK;
        const int rc = batch_complete(ctx, B, si);
        if (rc) return rc;
    }
}

static int batch_drain(topk_ctx *ctx, BatchState &B) { return batch_wait_upto(ctx, B, ~0ull); }

// Validates, stages and launches one batch (both passes in one launch) in
// the next slot; a busy slot holds the older of the two batches in flight
// and is completed first.
// Real-data scores (divided by qq after every term, topk_set_tuning
// "score_divide"; the reference's LD == 2 findTopKsimdMQ, cpu/TPAc.h:237-240)
// have no partition bounds here: every query of the batch is answered by
// the FTE path, synchronously, and its ticket is complete on return.
static int batch_submit_divided(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t ks, uint64_t *out_ids,
                                float *out_scores, float *out_thr, uint64_t *ticket) {
    const double t0 = now_ms();
    if (!ctx->batch) ctx->batch = new BatchState();
    BatchState &B = *static_cast<BatchState *>(ctx->batch);
    int rc = batch_drain(ctx, B);
    if (rc) return rc;
    for (uint32_t i = 0; i < nq; i++) {
        const topk_query &Q = qs[i];
        if (Q.k < 1 || Q.k > ks || Q.qq < 1 || Q.qq > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
        if (Q.k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    }
    uint64_t candidates = 0;
    for (uint32_t i = 0; i < nq; i++) {
        const topk_query &Q = qs[i];
        float t = 0;
        rc = topk_fte(ctx, Q.k, (uint8_t)Q.qq, Q.weights, Q.attr, out_ids ? out_ids + (uint64_t)i * ks : nullptr,
                      out_scores ? out_scores + (uint64_t)i * ks : nullptr, &t);
        if (rc) return rc;
        if (out_thr) out_thr[i] = t;
        candidates += ctx->stats.candidates;
    }
    topk_stats &S = ctx->stats;
    S = topk_stats{};
    S.queries = nq;
    S.fte_fallback = nq;
    S.candidates = candidates;
    S.tuple_count = (uint64_t)nq * ctx->n;
    for (uint32_t i = 0; i < nq; i++) S.terms += ctx->n * qs[i].qq;
    S.ms_last = now_ms() - t0;
    if (ticket) *ticket = B.next_ticket;
    B.next_ticket++;
    return TOPK_OK;
}

static int batch_submit(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t ks, uint64_t *out_ids,
                        float *out_scores, float *out_thr, const BatchVariant *force, uint64_t *ticket) {
    if (!ctx || !qs || nq == 0 || ks == 0) return TOPK_ERR_INVALID;
    if (nq > BATCH_MAX_Q || ks > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    // real-data scores, and stores with 64-bit ids (no partition layout):
    // every query through the FTE path
    if (ctx->score_div || ctx->n >= ID_LIMIT) {
        if (force) return TOPK_ERR_UNSUPPORTED;  // no launch shapes to time
        return batch_submit_divided(ctx, nq, qs, ks, out_ids, out_scores, out_thr, ticket);
    }
    const double t0 = now_ms();
    if (!ctx->pta.built) {
        const int rc = pta_build(ctx);
        if (rc) return rc;
    }
    if (!ctx->batch) ctx->batch = new BatchState();
    BatchState &B = *static_cast<BatchState *>(ctx->batch);
    const int mode = resolve_mode(ctx);
    if (mode < 0) return TOPK_ERR_STATE;
    const int si = B.next;
    int rc = batch_complete(ctx, B, si);
    if (!rc) rc = ensure(ctx, B, nq, ks, si);
    if (rc) return rc;
    const BatchVariant v = force ? *force : pick_variant(ctx, B);  // after ensure: the CU count is known
    if (!variant_ok(dims_class(ctx->d), v.vpt, v.bs) || v.qpi < 16 || v.qpi > 1024 ||
        (v.digit != 8 && v.digit != 11) || v.grid < 1 || v.grid > (int)BATCH_MAX_GRID || v.item < (int)PTA_CHUNK ||
        v.item % PTA_CHUNK)
        return TOPK_ERR_INVALID;
    B.cur = si;
    rc = stage_queries(ctx, B, nq, qs, ks);
    if (rc) return rc;
    HostSlot &H = B.slot[si];
    hipStream_t s = H.stream;
    // Every query filling its stride (k = k_stride <= n) into pinned caller
    // arrays: the batch's publish kernel writes them directly; otherwise it
    // fills the staging block and the wait copies out.
    bool full = ctx->n >= ks;
    for (uint32_t i = 0; full && i < nq; i++) full = qs[i].k == ks;
    H.direct = full && pinned_host(B, out_ids) && pinned_host(B, out_scores) && pinned_host(B, out_thr);
    OutDesc od{};
    if (H.direct) {
        od.ids = reinterpret_cast<uint64_t>(out_ids);
        od.scores = reinterpret_cast<uint64_t>(out_scores);
        od.thr = reinterpret_cast<uint64_t>(out_thr);
        od.direct = 1;
    }
    memcpy(H.h_in + InLayout::make(nq).od_off, &od, sizeof(od));
    // upload, kernels and publish on the slot's stream (one graph launch)
    rc = batch_launch(ctx, B, s, nq, ks, mode, v);
    if (rc) return rc;
    TK_HIP(hipEventRecord(H.done, s));
    H.busy = true;
    H.ticket = B.next_ticket++;
    H.nq = nq;
    H.ks = ks;
    H.out_ids = out_ids;
    H.out_scores = out_scores;
    H.out_thr = out_thr;
    H.t0 = t0;
    B.next = si ^ 1;
    if (ticket) *ticket = H.ticket;
    return TOPK_OK;
}

static int batch_run(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t ks, uint64_t *out_ids,
                     float *out_scores, float *out_thr, const BatchVariant *force) {
    uint64_t t = 0;
    const int rc = batch_submit(ctx, nq, qs, ks, out_ids, out_scores, out_thr, force, &t);
    if (rc) return rc;
    return batch_wait_upto(ctx, *static_cast<BatchState *>(ctx->batch), t);
}

void batch_quiesce(topk_ctx *ctx) {
    if (ctx && ctx->batch) batch_drain(ctx, *static_cast<BatchState *>(ctx->batch));
}

}  // namespace topk

using namespace topk;

extern "C" {

int topk_batch(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride, uint64_t *out_ids,
               float *out_scores, float *out_thresholds) {
    return batch_run(ctx, nq, qs, k_stride, out_ids, out_scores, out_thresholds, nullptr);
}

int topk_batch_submit(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride, uint64_t *out_ids,
                      float *out_scores, float *out_thresholds, uint64_t *ticket) {
    return batch_submit(ctx, nq, qs, k_stride, out_ids, out_scores, out_thresholds, nullptr, ticket);
}

int topk_batch_wait(topk_ctx *ctx, uint64_t ticket) {
    if (!ctx) return TOPK_ERR_INVALID;
    if (!ctx->batch) return ticket ? TOPK_ERR_INVALID : TOPK_OK;
    BatchState &B = *static_cast<BatchState *>(ctx->batch);
    if (ticket >= B.next_ticket) return TOPK_ERR_INVALID;
    return batch_wait_upto(ctx, B, ticket);
}

int topk_batch_autotune(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride,
                        topk_batch_tuning *out) {
    if (!ctx || !qs || nq == 0 || k_stride == 0) return TOPK_ERR_INVALID;
    if (nq > BATCH_MAX_Q || k_stride > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    if (ctx->score_div || ctx->n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;  // those batches run the FTE path
    if (!ctx->pta.built) {
        const int rc = pta_build(ctx);
        if (rc) return rc;
    }
    if (!ctx->batch) ctx->batch = new BatchState();
    BatchState &B = *static_cast<BatchState *>(ctx->batch);
    const int mode = resolve_mode(ctx);
    if (mode < 0) return TOPK_ERR_STATE;
    int rc = batch_drain(ctx, B);
    if (!rc) rc = ensure(ctx, B, nq, k_stride, 0);
    if (rc) return rc;
    B.cur = 0;
    rc = stage_queries(ctx, B, nq, qs, k_stride);
    if (rc) return rc;
    hipStream_t s = B.slot[0].stream;
    {
        const OutDesc od{};  // results to the staging block
        memcpy(B.slot[0].h_in + InLayout::make(nq).od_off, &od, sizeof(od));
    }
    const int D = dims_class(ctx->d);
    const BatchVariant def = pick_variant(ctx, B);
    std::vector<BatchVariant> cand;
    static const int vpts[] = {2, 4, 8, 16}, qpis[] = {16, 32, 64, 128}, per_cu[] = {2, 4, 8};
    for (int vpt : vpts)
        for (int bs : {256, 512, 1024}) {
            if (!variant_ok(D, vpt, bs)) continue;
            for (int qpi : qpis)
                for (int pc : per_cu)
                    cand.push_back(BatchVariant{vpt, bs, qpi, def.digit, B.cus * pc * 256 / bs, def.item, 0});
        }
    hipEvent_t e0, e1;
    TK_HIP(hipEventCreate(&e0));
    TK_HIP(hipEventCreate(&e1));
    auto time_one = [&](const BatchVariant &v, double &ms) -> int {
        ms = 1e30;
        for (int rep = 0; rep < 3; rep++) {
            TK_HIP(hipEventRecord(e0, s));
            const int r = batch_enqueue(ctx, B, s, nq, k_stride, mode, v);
            if (r) return r;
            TK_HIP(hipEventRecord(e1, s));
            TK_HIP(hipEventSynchronize(e1));
            float t = 0;
            TK_HIP(hipEventElapsedTime(&t, e0, e1));
            if (rep && t < ms) ms = t;  // first run warms the code and caches
        }
        return TOPK_OK;
    };
    {
        BatchVariant f = def;  // the query-major fused pass (its shape knobs do not apply)
        f.fused = 1;
        cand.push_back(f);
    }
    double ms_def = 0, best_ms = 1e30;
    BatchVariant best = def;
    rc = time_one(def, ms_def);
    best_ms = ms_def;
    for (size_t i = 0; !rc && i < cand.size(); i++) {
        double ms;
        rc = time_one(cand[i], ms);
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best = cand[i];
        }
    }
    uint32_t extra = 0;
    for (int dg : {8, 11}) {  // the final select's digit width, with the best shape
        if (rc || dg == best.digit) continue;
        BatchVariant v = best;
        v.digit = dg;
        double ms;
        rc = time_one(v, ms);
        extra++;
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    for (int it : {2048, 4096, 8192, 16384}) {  // objects per work item, with the best shape
        if (rc || it == best.item || best.fused) continue;
        BatchVariant v = best;
        v.item = it;
        double ms;
        rc = time_one(v, ms);
        extra++;
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    ctx->tune.batch_vpt = best.vpt;
    ctx->tune.batch_bs = best.bs;
    ctx->tune.batch_qpi = best.qpi;
    ctx->tune.batch_digit = best.digit;
    ctx->tune.batch_grid = best.grid;
    ctx->tune.batch_item = best.item;
    ctx->tune.batch_fused = best.fused ? 2 : 1;
    if (out) {
        out->vpt = (uint32_t)best.vpt;
        out->bs = (uint32_t)best.bs;
        out->qpi = (uint32_t)best.qpi;
        out->digit = (uint32_t)best.digit;
        out->grid = (uint32_t)best.grid;
        out->item = (uint32_t)best.item;
        out->variants = (uint32_t)cand.size() + 1 + extra;
        out->ms_best = best_ms;
        out->ms_default = ms_def;
        out->fused = (uint32_t)best.fused;
    }
    return TOPK_OK;
}

}  // extern "C"
