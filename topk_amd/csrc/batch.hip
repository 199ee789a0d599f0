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
//  1. threshold per query, tau_q (a score: objects >= tau_q are candidates)
//       - seed mode: LP bounds of every (query, partition) pair; the exact
//         k-th score among the first objects of the query's 4 best-bounded
//         partitions is a proven lower bound of the query's k-th score.
//       - learned mode: an MLP predicts the k-th score from (query, data
//         sketch), lowered by a calibrated margin (learned.hip).
//  2. mask: pair (p, q) is scored iff bound(p, q) >= tau_q; set bits are
//     OR-ed into a per-partition query bitmap.
//  3. compaction: non-empty (partition, query group) bitmaps become work
//     items.
//  4. scoring (persistent workgroups taking items from a global ticket):
//     a workgroup loads the partition's objects into registers once and
//     scores them for every query of the item; survivors (score >= tau_q)
//     are appended to the query's candidate list (wave-aggregated atomic).
//     Scores use the query's own term order: a wave-uniform switch on the
//     attribute id picks the register column, so any attribute order,
//     subset or repeat costs the same as the plain weighted sum.
//  5. final select per query (one workgroup each): radix select in LDS over
//     the candidates held in registers, rank sort, best-first output.
//
// Exactness: every object with score >= tau_q lies in a partition whose
// bound is >= tau_q, so it is a candidate.  With c >= k candidates the top k
// are among them.  A learned tau_q above the true k-th score leaves c < k
// (status RETRY) and a candidate list over its capacity raises OVERFLOW;
// both queries are answered by the plain exact FTE path on the host side of
// the call, and counted in topk_stats.fallback.
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
constexpr int BFIN_BS = 512, BFIN_KPT = 16;  // lists up to 8192 keys are selected in registers, longer ones in memory
constexpr int BSEED_BS = 512, BSEED_KPT = 32;
constexpr uint32_t BSEED_T = 4;                          // seed partitions per query
constexpr uint32_t BSEED_OBJS = BSEED_BS * BSEED_KPT;    // objects a seed scores (16384)
constexpr int BAUX_BS = 256;                             // bound / mask / compaction kernels
enum { BMODE_SEED = 1, BMODE_LEARNED = 2, BMODE_GIVEN = 3 };
// Device counters of one batch: [0] work items, [1] work ticket,
// [2..3] scored (partition, query) pairs, [4..5] scored (object, query) pairs.
constexpr uint32_t BCTR_WORDS = 8;

struct BatchItem {
    uint32_t p;                 // partition
    uint32_t q0;                // query of bit 0
    unsigned long long mask;    // queries q0 + b for the set bits b
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BAUX_BS) void batch_zero_kernel(unsigned long long *__restrict__ masks, uint64_t nmask,
                                                             unsigned int *__restrict__ ccount, uint32_t nq,
                                                             unsigned int *__restrict__ ctr) {
    const uint64_t stride = (uint64_t)gridDim.x * BAUX_BS;
    const uint64_t i0 = (uint64_t)blockIdx.x * BAUX_BS + threadIdx.x;
    for (uint64_t i = i0; i < nmask; i += stride) masks[i] = 0ull;
    for (uint64_t i = i0; i < nq; i += stride) ccount[i] = 0u;
    if (i0 < BCTR_WORDS) ctr[i0] = 0u;
}

// bnd[q * L + p] = LP bound of query q over partition p (-inf when empty).
__global__ __launch_bounds__(BAUX_BS) void batch_bounds_kernel(const BoundArgs *__restrict__ qb, uint32_t L,
                                                               const uint32_t *__restrict__ psz,
                                                               const float *__restrict__ pmax,
                                                               const float *__restrict__ pmin,
                                                               const float *__restrict__ psum,
                                                               const float *__restrict__ pminsum,
                                                               float *__restrict__ bnd) {
    const uint32_t p = blockIdx.x * BAUX_BS + threadIdx.x, q = blockIdx.y;
    if (p >= L) return;
    const BoundArgs &b = qb[q];
    float ub = -INFINITY;
    if (psz[p]) ub = lp_bound(b, pmax + (uint64_t)p * b.d, pmin + (uint64_t)p * b.d, psum[p], pminsum[p]);
    bnd[(uint64_t)q * L + p] = ub;
}

// Bound keys of one query's partitions (0 for empty partitions), uniform
// trip count so the gather's ballots see every lane.
template <int BS>
struct BoundKeys {
    const float *b;
    const uint32_t *psz;
    uint32_t L;
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
        for (uint32_t i0 = 0; i0 < L; i0 += BS) {
            const uint32_t i = i0 + threadIdx.x;
            f(i < L && psz[i] ? rkey(b[i], i) : 0ull);
        }
    }
};

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

// Seed pass, one workgroup per query: the 4 partitions with the largest
// bound keys, the first objects of each (16384 in all), and the exact k-th
// largest key among them; its score is tau.  Fewer than k objects: tau =
// -inf (every object is a candidate).
__global__ __launch_bounds__(BSEED_BS) void batch_seed_kernel(BatchQuery *__restrict__ qs, const float *__restrict__ bnd,
                                                              uint32_t L, uint32_t nonempty,
                                                              const float *__restrict__ pcols, uint64_t ld,
                                                              const uint32_t *__restrict__ poff,
                                                              const uint32_t *__restrict__ psz,
                                                              const uint32_t *__restrict__ pid) {
    __shared__ SelectScratch ss;
    __shared__ uint64_t best[BSEED_T];
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const BatchQuery &Q = qs[q];
    const BoundKeys<BSEED_BS> src{bnd + (uint64_t)q * L, psz, L};
    const uint32_t T = nonempty < BSEED_T ? nonempty : BSEED_T;
    const uint64_t kth = block_kth_largest<BSEED_BS>(src, nonempty, T, ss);
    block_gather_ge<BSEED_BS>(src, kth, best, ss);  // exactly T keys (keys are distinct)
    const uint32_t per = BSEED_OBJS / T;
    uint32_t start[BSEED_T], lim[BSEED_T], total = 0;
#pragma unroll
    for (uint32_t i = 0; i < BSEED_T; i++) {
        const uint32_t p = i < T ? rkey_id(best[i]) : 0u;
        start[i] = i < T ? poff[p] : 0u;
        lim[i] = i < T ? (psz[p] < per ? psz[p] : per) : 0u;
        total += lim[i];
    }
    RegKeys<BSEED_KPT> rk;
#pragma unroll
    for (int j = 0; j < BSEED_KPT; j++) {
        const uint32_t g = j * BSEED_BS + tid;
        const uint32_t i = g / per, o = g - i * per;
        uint64_t key = 0;
        if (i < T && o < lim[i]) {
            const uint64_t pos = (uint64_t)start[i] + o;
            key = rkey(batch_score_one(Q, pcols, ld, pos), pid[pos]);
        }
        rk.k[j] = key;
    }
    float tau = -INFINITY;
    if (total >= Q.k) tau = rkey_score(block_kth_largest<BSEED_BS>(rk, total, Q.k, ss));
    if (tid == 0) qs[q].tau = tau;
}

// Sets bit q of masks[p * nw + q / 64] for every pair with bound >= tau_q.
template <bool STORED>
__global__ __launch_bounds__(BAUX_BS) void batch_mask_kernel(const BatchQuery *__restrict__ qs,
                                                             const BoundArgs *__restrict__ qb,
                                                             const float *__restrict__ bnd, uint32_t L, uint32_t nw,
                                                             const uint32_t *__restrict__ psz,
                                                             const float *__restrict__ pmax,
                                                             const float *__restrict__ pmin,
                                                             const float *__restrict__ psum,
                                                             const float *__restrict__ pminsum,
                                                             unsigned long long *__restrict__ masks) {
    const uint32_t p = blockIdx.x * BAUX_BS + threadIdx.x, q = blockIdx.y;
    if (p >= L || !psz[p]) return;
    const float tau = qs[q].tau;
    float ub;
    if constexpr (STORED) {
        ub = bnd[(uint64_t)q * L + p];
    } else {
        const BoundArgs &b = qb[q];
        ub = lp_bound(b, pmax + (uint64_t)p * b.d, pmin + (uint64_t)p * b.d, psum[p], pminsum[p]);
    }
    if (ub >= tau) atomicOr(&masks[(uint64_t)p * nw + q / 64], 1ull << (q % 64));
}

// Non-empty groups of qpi queries of every partition bitmap become items
// (one global atomic per workgroup), with the pair counters updated.
__global__ __launch_bounds__(BAUX_BS) void batch_compact_kernel(const unsigned long long *__restrict__ masks,
                                                                uint64_t nmask, uint32_t nw, uint32_t qpi,
                                                                const uint32_t *__restrict__ psz,
                                                                BatchItem *__restrict__ items,
                                                                unsigned int *__restrict__ ctr) {
    __shared__ uint32_t wsum[BAUX_BS / 64];
    __shared__ uint32_t total, base;
    const uint32_t nch = 64 / qpi;
    const uint64_t t = (uint64_t)blockIdx.x * BAUX_BS + threadIdx.x;
    unsigned long long sub = 0;
    uint64_t pw = 0;
    uint32_t c = 0;
    if (t < nmask * nch) {
        pw = t / nch;
        c = (uint32_t)(t - pw * nch);
        const unsigned long long cm = qpi == 64 ? ~0ull : (((1ull << qpi) - 1ull) << (c * qpi));
        sub = masks[pw] & cm;
    }
    const uint32_t flag = sub ? 1u : 0u;
    const uint32_t pos = block_excl_scan<BAUX_BS>(flag, wsum, &total);
    const uint32_t np = (uint32_t)__popcll(sub);
    const uint32_t pairs = wave_sum_u32(np);
    unsigned long long objs = (unsigned long long)np * (sub ? psz[pw / nw] : 0u);
    for (int o = 32; o > 0; o >>= 1) objs += __shfl_xor(objs, o, 64);
    if (threadIdx.x == 0) base = total ? atomicAdd(&ctr[0], total) : 0u;
    if (lane_id() == 0 && pairs) {
        atomicAdd(reinterpret_cast<unsigned long long *>(ctr + 2), (unsigned long long)pairs);
        atomicAdd(reinterpret_cast<unsigned long long *>(ctr + 4), objs);
    }
    __syncthreads();
    if (flag) {
        BatchItem it;
        it.p = (uint32_t)(pw / nw);
        it.q0 = (uint32_t)(pw % nw) * 64u;
        it.mask = sub;
        items[base + pos] = it;
    }
}

struct ScoreArgs {
    const float *pcols;
    uint64_t ld;
    uint32_t d;
    const uint32_t *poff, *psz, *pid;
    const BatchItem *items;
    unsigned int *ctr;
    const BatchQuery *q;
    uint64_t *cand;
    unsigned int *ccount;
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
template <int D, int VPT, int BS>
__global__ __launch_bounds__(BS) void batch_score_kernel(ScoreArgs a) {
    __shared__ uint32_t s_it;
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t nitems = __builtin_amdgcn_readfirstlane(a.ctr[0]);
    for (;;) {
        if (tid == 0) s_it = atomicAdd(&a.ctr[1], 1u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_it);
        __syncthreads();
        if (it >= nitems) break;
        const BatchItem I = a.items[it];
        const uint32_t start = a.poff[I.p], size = a.psz[I.p];
        for (uint32_t b0 = 0; b0 < size; b0 += BS * VPT) {
            const uint32_t i0 = b0 + tid * VPT;
            const uint32_t nvalid = i0 < size ? (size - i0 < (uint32_t)VPT ? size - i0 : (uint32_t)VPT) : 0u;
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
            unsigned long long mask = I.mask;
            while (mask) {
                const uint32_t bit = (uint32_t)__builtin_ctzll(mask);
                mask &= mask - 1ull;
                const uint32_t q = I.q0 + bit;
                const BatchQuery &Q = a.q[q];
                float s[VPT];
#pragma unroll
                for (int e = 0; e < VPT; e++) s[e] = 0.0f;
                for (uint32_t j = 0; j < Q.m; j++) {
                    const float w = Q.w[j];
                    switch (Q.attr[j]) {
                        TK_TERM(0) TK_TERM(1) TK_TERM(2) TK_TERM(3) TK_TERM(4) TK_TERM(5) TK_TERM(6) TK_TERM(7)
                        TK_TERM(8) TK_TERM(9) TK_TERM(10) TK_TERM(11) TK_TERM(12) TK_TERM(13) TK_TERM(14)
                        TK_TERM(15) TK_TERM(16) TK_TERM(17) TK_TERM(18) TK_TERM(19) TK_TERM(20) TK_TERM(21)
                        TK_TERM(22) TK_TERM(23) TK_TERM(24) TK_TERM(25) TK_TERM(26) TK_TERM(27) TK_TERM(28)
                        TK_TERM(29) TK_TERM(30) TK_TERM(31)
                        default: break;
                    }
                }
                const float tau = Q.tau;
                uint32_t np = 0;
#pragma unroll
                for (int e = 0; e < VPT; e++) np += ((uint32_t)e < nvalid && s[e] >= tau) ? 1u : 0u;
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
                        if ((uint32_t)e < nvalid && s[e] >= tau) {
                            if (pos < BATCH_CAP) out[pos] = rkey(s[e], a.pid[start + i0 + e]);
                            pos++;
                        }
                    }
                }
            }
        }
    }
}
#undef TK_TERM

struct FinalArgs {
    const BatchQuery *q;
    const uint64_t *cand;
    const unsigned int *ccount;
    const unsigned int *ctr;
    uint32_t kstride;
    uint64_t *ids;
    float *scores;
    float *thr;
    uint32_t *status;
    uint32_t *cnt;
    uint32_t *ctr_out;
    float *tau_next;
};

template <int DB>
__global__ __launch_bounds__(BFIN_BS) void batch_final_kernel(FinalArgs a) {
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ uint64_t sorted[TOPK_MAX_K];
    __shared__ SelectScratch ss;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t c = a.ccount[q], k = a.q[q].k;
    if (q == 0 && tid < BCTR_WORDS) a.ctr_out[tid] = a.ctr[tid];
    if (tid == 0) a.cnt[q] = c;
    if (c > BATCH_CAP) {
        // The stored keys are a subset of the objects, so their k-th largest
        // is a lower bound of the query's k-th key: a tighter threshold for a
        // second pass.
        const uint64_t kth = block_kth_largest<BFIN_BS>(MemKeys<BFIN_BS>{a.cand + (uint64_t)q * BATCH_CAP, BATCH_CAP},
                                                        BATCH_CAP, k, ss);
        if (tid == 0) {
            a.status[q] = STATUS_OVERFLOW;
            a.tau_next[q] = rkey_score(kth);
        }
        return;
    }
    if (c < k) {
        if (tid == 0) a.status[q] = STATUS_RETRY;
        return;
    }
    const FinalOut o{a.ids + (uint64_t)q * a.kstride, a.scores + (uint64_t)q * a.kstride, a.thr + q, a.status + q};
    select_and_emit<BFIN_BS, BFIN_KPT, FlatView, DB>(FlatView{a.cand + (uint64_t)q * BATCH_CAP}, c, k, o, top, sorted,
                                                     ss);
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
struct BatchVariant {
    int vpt, bs, qpi, digit, grid;
};

using GraphKey = std::array<uint64_t, 10>;

struct BatchState {
    uint64_t gen = ~0ull;   // layout generation the buffers and graphs belong to
    uint32_t nq_cap = 0;
    size_t in_cap = 0, out_cap = 0;
    uint64_t mask_cap = 0, item_cap = 0, bnd_cap = 0;
    unsigned char *d_in = nullptr, *h_in = nullptr;    // BatchQuery[nq] | BoundArgs[nq]
    unsigned char *d_out = nullptr, *h_out = nullptr;  // ids | scores | thr | status | cnt | ctr
    unsigned long long *d_masks = nullptr;
    BatchItem *d_items = nullptr;
    unsigned int *d_ctr = nullptr;
    unsigned int *d_ccount = nullptr;
    uint64_t *d_cand = nullptr;
    float *d_bnd = nullptr;
    int cus = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
};

struct InLayout {
    size_t q_off, b_off, bytes;
    static InLayout make(uint32_t nq) {
        InLayout l;
        l.q_off = 0;
        l.b_off = (sizeof(BatchQuery) * nq + 255) & ~(size_t)255;
        l.bytes = l.b_off + sizeof(BoundArgs) * nq;
        return l;
    }
};

struct OutLayout {
    size_t ids_off, sc_off, thr_off, st_off, cnt_off, tn_off, ctr_off, bytes;
    static OutLayout make(uint32_t nq, uint32_t ks) {
        OutLayout l;
        l.ids_off = 0;
        l.sc_off = l.ids_off + sizeof(uint64_t) * nq * ks;
        l.thr_off = l.sc_off + ((sizeof(float) * nq * ks + 15) & ~(size_t)15);
        l.st_off = l.thr_off + sizeof(float) * nq;
        l.cnt_off = l.st_off + sizeof(uint32_t) * nq;
        l.tn_off = l.cnt_off + sizeof(uint32_t) * nq;
        l.ctr_off = (l.tn_off + sizeof(float) * nq + 15) & ~(size_t)15;
        l.bytes = l.ctr_off + sizeof(uint32_t) * BCTR_WORDS;
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
    drop_graphs(*B);
    hipFree(B->d_in);
    hipFree(B->d_out);
    hipFree(B->d_masks);
    hipFree(B->d_items);
    hipFree(B->d_ctr);
    hipFree(B->d_ccount);
    hipFree(B->d_cand);
    hipFree(B->d_bnd);
    if (B->h_in) hipHostFree(B->h_in);
    if (B->h_out) hipHostFree(B->h_out);
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

static int ensure(topk_ctx *ctx, BatchState &B, uint32_t nq, uint32_t ks, bool need_bnd) {
    const PtaLayout &P = ctx->pta;
    bool changed = false;
    if (B.gen != P.generation) {
        changed = true;
        B.gen = P.generation;
    }
    const uint32_t L = P.nparts, nw = (nq + 63) / 64;
    if (nq > B.nq_cap) {
        uint32_t cap = 64;
        while (cap < nq) cap *= 2;
        hipFree(B.d_ccount);
        hipFree(B.d_cand);
        B.d_ccount = nullptr;
        B.d_cand = nullptr;
        B.nq_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&B.d_ccount, sizeof(unsigned int) * cap));
        TK_HIP(hipMalloc(&B.d_cand, sizeof(uint64_t) * BATCH_CAP * cap));
        B.nq_cap = cap;
    }
    const size_t in_need = InLayout::make(nq).bytes, out_need = OutLayout::make(nq, ks).bytes;
    if (in_need > B.in_cap) {
        hipFree(B.d_in);
        if (B.h_in) hipHostFree(B.h_in);
        B.d_in = B.h_in = nullptr;
        B.in_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&B.d_in, in_need));
        TK_HIP(hipHostMalloc(&B.h_in, in_need, hipHostMallocDefault));
        B.in_cap = in_need;
    }
    if (out_need > B.out_cap) {
        hipFree(B.d_out);
        if (B.h_out) hipHostFree(B.h_out);
        B.d_out = B.h_out = nullptr;
        B.out_cap = 0;
        changed = true;
        TK_HIP(hipMalloc(&B.d_out, out_need));
        TK_HIP(hipHostMalloc(&B.h_out, out_need, hipHostMallocDefault));
        B.out_cap = out_need;
    }
    int rc = grow_dev(B.d_masks, B.mask_cap, (uint64_t)L * nw, changed);
    if (!rc) rc = grow_dev(B.d_items, B.item_cap, (uint64_t)L * nw * 4, changed);  // qpi >= 16: 4 items per word
    if (!rc && need_bnd) rc = grow_dev(B.d_bnd, B.bnd_cap, (uint64_t)L * nq, changed);
    if (rc) return rc;
    if (!B.d_ctr) {
        changed = true;
        TK_HIP(hipMalloc(&B.d_ctr, sizeof(unsigned int) * BCTR_WORDS));
    }
    if (!B.cus) {
        TK_HIP(hipDeviceGetAttribute(&B.cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (B.cus < 1) B.cus = 1;
    }
    if (changed) drop_graphs(B);
    return TOPK_OK;
}

static int dims_class(uint32_t d) { return d <= 8 ? 8 : (d <= 16 ? 16 : 32); }

static BatchVariant pick_variant(const topk_ctx *ctx, const BatchState &B) {
    const int D = dims_class(ctx->d);
    BatchVariant v;
    v.vpt = ctx->tune.batch_vpt ? ctx->tune.batch_vpt : (D == 8 ? 8 : (D == 16 ? 4 : 2));
    v.bs = ctx->tune.batch_bs ? ctx->tune.batch_bs : 256;
    v.qpi = ctx->tune.batch_qpi ? ctx->tune.batch_qpi : 64;
    v.digit = ctx->tune.batch_digit ? ctx->tune.batch_digit : 11;
    v.grid = ctx->tune.batch_grid ? ctx->tune.batch_grid : B.cus * (v.bs == 256 ? 4 : 2);
    return v;
}

static bool variant_ok(int D, int vpt, int bs) {
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
    TK_SC(16, 2, 256) TK_SC(16, 4, 256) TK_SC(16, 8, 256) TK_SC(16, 2, 512) TK_SC(16, 4, 512) TK_SC(16, 8, 512)
    TK_SC(32, 2, 256) TK_SC(32, 4, 256) TK_SC(32, 2, 512) TK_SC(32, 4, 512)
#undef TK_SC
}

// Enqueues the whole batch pipeline (staged inputs already in h_in) on s.
static int batch_enqueue(topk_ctx *ctx, BatchState &B, hipStream_t s, uint32_t nq, uint32_t ks, int mode,
                         const BatchVariant &v) {
    const PtaLayout &P = ctx->pta;
    const uint32_t L = P.nparts, nw = (nq + 63) / 64;
    const uint64_t nmask = (uint64_t)L * nw;
    const InLayout il = InLayout::make(nq);
    const OutLayout ol = OutLayout::make(nq, ks);
    BatchQuery *dq = reinterpret_cast<BatchQuery *>(B.d_in + il.q_off);
    const BoundArgs *db = reinterpret_cast<const BoundArgs *>(B.d_in + il.b_off);
    TK_HIP(hipMemcpyAsync(B.d_in, B.h_in, il.bytes, hipMemcpyHostToDevice, s));
    uint32_t zg = (uint32_t)((nmask + BAUX_BS - 1) / BAUX_BS);
    if (zg > 2048) zg = 2048;
    if (zg < 1) zg = 1;
    hipLaunchKernelGGL(batch_zero_kernel, dim3(zg), dim3(BAUX_BS), 0, s, B.d_masks, nmask, B.d_ccount, nq, B.d_ctr);
    const dim3 pq((L + BAUX_BS - 1) / BAUX_BS, nq);
    if (mode == BMODE_SEED) {
        hipLaunchKernelGGL(batch_bounds_kernel, pq, dim3(BAUX_BS), 0, s, db, L, P.psz, P.pmax, P.pmin, P.psum,
                           P.pminsum, B.d_bnd);
        hipLaunchKernelGGL(batch_seed_kernel, dim3(nq), dim3(BSEED_BS), 0, s, dq, B.d_bnd, L, P.nonempty, P.pcols,
                           P.ld, P.poff, P.psz, P.pid);
        hipLaunchKernelGGL(batch_mask_kernel<true>, pq, dim3(BAUX_BS), 0, s, dq, db, B.d_bnd, L, nw, P.psz, P.pmax,
                           P.pmin, P.psum, P.pminsum, B.d_masks);
    } else {
        if (mode == BMODE_LEARNED) {
            const int rc = learned_tau_enqueue(ctx, s, dq, nq);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(batch_mask_kernel<false>, pq, dim3(BAUX_BS), 0, s, dq, db, nullptr, L, nw, P.psz, P.pmax,
                           P.pmin, P.psum, P.pminsum, B.d_masks);
    }
    const uint64_t nthreads = nmask * (64 / v.qpi);
    hipLaunchKernelGGL(batch_compact_kernel, dim3((uint32_t)((nthreads + BAUX_BS - 1) / BAUX_BS)), dim3(BAUX_BS), 0,
                       s, B.d_masks, nmask, nw, (uint32_t)v.qpi, P.psz, B.d_items, B.d_ctr);
    ScoreArgs sa{P.pcols, P.ld, ctx->d, P.poff, P.psz, P.pid, B.d_items, B.d_ctr, dq, B.d_cand, B.d_ccount};
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    launch_score(dims_class(ctx->d), v, s, sa);
    timing_end(ctx, s, "batch_score", ev0);
    FinalArgs fa{dq,
                 B.d_cand,
                 B.d_ccount,
                 B.d_ctr,
                 ks,
                 reinterpret_cast<uint64_t *>(B.d_out + ol.ids_off),
                 reinterpret_cast<float *>(B.d_out + ol.sc_off),
                 reinterpret_cast<float *>(B.d_out + ol.thr_off),
                 reinterpret_cast<uint32_t *>(B.d_out + ol.st_off),
                 reinterpret_cast<uint32_t *>(B.d_out + ol.cnt_off),
                 reinterpret_cast<uint32_t *>(B.d_out + ol.ctr_off),
                 reinterpret_cast<float *>(B.d_out + ol.tn_off)};
    if (v.digit == 8)
        hipLaunchKernelGGL(batch_final_kernel<8>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
    else
        hipLaunchKernelGGL(batch_final_kernel<11>, dim3(nq), dim3(BFIN_BS), 0, s, fa);
    TK_HIP(hipGetLastError());
    TK_HIP(hipMemcpyAsync(B.h_out, B.d_out, ol.bytes, hipMemcpyDeviceToHost, s));
    return TOPK_OK;
}

static int batch_launch(topk_ctx *ctx, BatchState &B, hipStream_t s, uint32_t nq, uint32_t ks, int mode,
                        const BatchVariant &v) {
    if (!ctx->tune.batch_graph || ctx->timing_on) return batch_enqueue(ctx, B, s, nq, ks, mode, v);
    const GraphKey key{nq, ks, (uint64_t)mode, (uint64_t)v.vpt, (uint64_t)v.bs, (uint64_t)v.qpi, (uint64_t)v.digit,
                       (uint64_t)v.grid, ctx->d, mode == BMODE_LEARNED ? learned_version(ctx) : 0};
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
    for (uint32_t j = 0; j < Q.qq; j++) {
        if (Q.attr[j] >= ctx->d) return TOPK_ERR_INVALID;
        b.attr[j] = Q.attr[j];
        b.w[j] = Q.weights[Q.attr[j]];
    }
    return TOPK_OK;
}

namespace topk {

// Validates and stages the queries into h_in.
static int stage_queries(topk_ctx *ctx, BatchState &B, uint32_t nq, const topk_query *qs, uint32_t ks) {
    const InLayout il = InLayout::make(nq);
    BatchQuery *hq = reinterpret_cast<BatchQuery *>(B.h_in + il.q_off);
    BoundArgs *hb = reinterpret_cast<BoundArgs *>(B.h_in + il.b_off);
    for (uint32_t i = 0; i < nq; i++) {
        const topk_query &Q = qs[i];
        const int rc = make_batch_query(ctx, Q, hq[i]);
        if (rc) return rc;
        if (Q.k > ks) return TOPK_ERR_INVALID;
        hb[i] = make_bound_args(ctx->d, Q.qq, Q.weights, Q.attr);
    }
    return TOPK_OK;
}

static int resolve_mode(const topk_ctx *ctx) {
    if (ctx->tune.batch_mode == BMODE_SEED) return BMODE_SEED;
    if (ctx->tune.batch_mode == BMODE_LEARNED) return learned_ready(ctx) ? BMODE_LEARNED : -1;
    return learned_ready(ctx) ? BMODE_LEARNED : BMODE_SEED;
}

// Parsed outcome of one device pass.
struct PassTally {
    uint64_t pairs = 0, objs = 0, candidates = 0;
};

// Copies the finished queries of the last pass (local index i -> caller row
// rows[i]) and collects the others: overflowed queries with their tightened
// threshold, and queries with too few candidates.
static void collect(const topk_ctx *ctx, const BatchState &B, uint32_t nq, uint32_t ks, const topk_query *qs,
                    const uint32_t *rows, uint64_t *out_ids, float *out_scores, float *out_thr,
                    std::vector<std::pair<uint32_t, float>> &reseed, std::vector<uint32_t> &fte, PassTally &t) {
    const OutLayout ol = OutLayout::make(nq, ks);
    const uint64_t *ids = reinterpret_cast<const uint64_t *>(B.h_out + ol.ids_off);
    const float *sc = reinterpret_cast<const float *>(B.h_out + ol.sc_off);
    const float *thr = reinterpret_cast<const float *>(B.h_out + ol.thr_off);
    const uint32_t *st = reinterpret_cast<const uint32_t *>(B.h_out + ol.st_off);
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(B.h_out + ol.cnt_off);
    const float *tn = reinterpret_cast<const float *>(B.h_out + ol.tn_off);
    const uint32_t *ctr = reinterpret_cast<const uint32_t *>(B.h_out + ol.ctr_off);
    t.pairs += (uint64_t)ctr[2] | ((uint64_t)ctr[3] << 32);
    t.objs += (uint64_t)ctr[4] | ((uint64_t)ctr[5] << 32);
    for (uint32_t i = 0; i < nq; i++) {
        const uint32_t row = rows ? rows[i] : i;
        t.candidates += cnt[i] < BATCH_CAP ? cnt[i] : BATCH_CAP;
        if (st[i] == STATUS_OVERFLOW) {
            reseed.push_back({row, tn[i]});
            continue;
        }
        if (st[i]) {
            fte.push_back(row);
            continue;
        }
        const uint32_t kq = qs[row].k;
        const uint32_t c = (uint32_t)(ctx->n < kq ? ctx->n : kq);
        const uint64_t o = (uint64_t)row * ks, li = (uint64_t)i * ks;
        if (out_ids) {
            memcpy(out_ids + o, ids + li, sizeof(uint64_t) * c);
            for (uint32_t j = c; j < kq; j++) out_ids[o + j] = ~0ull;
        }
        if (out_scores) {
            memcpy(out_scores + o, sc + li, sizeof(float) * c);
            for (uint32_t j = c; j < kq; j++) out_scores[o + j] = -INFINITY;
        }
        if (out_thr) out_thr[row] = thr[i];
    }
}

// Runs one batch and copies results out.  Overflowed queries get a second
// device pass with the threshold their first pass tightened; queries the
// device path still cannot settle go through topk_fte.
static int batch_run(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t ks, uint64_t *out_ids,
                     float *out_scores, float *out_thr, const BatchVariant *force) {
    if (!ctx || !qs || nq == 0 || ks == 0) return TOPK_ERR_INVALID;
    if (nq > BATCH_MAX_Q || ks > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    const double t0 = now_ms();
    if (!ctx->pta.built) {
        const int rc = pta_build(ctx);
        if (rc) return rc;
    }
    if (!ctx->batch) ctx->batch = new BatchState();
    BatchState &B = *static_cast<BatchState *>(ctx->batch);
    const int mode = resolve_mode(ctx);
    if (mode < 0) return TOPK_ERR_STATE;
    int rc = ensure(ctx, B, nq, ks, mode == BMODE_SEED);
    if (rc) return rc;
    rc = stage_queries(ctx, B, nq, qs, ks);
    if (rc) return rc;
    const BatchVariant v = force ? *force : pick_variant(ctx, B);
    if (!variant_ok(dims_class(ctx->d), v.vpt, v.bs) || (v.qpi != 16 && v.qpi != 32 && v.qpi != 64) ||
        (v.digit != 8 && v.digit != 11) || v.grid < 1)
        return TOPK_ERR_INVALID;
    hipStream_t s = ctx->lane[0].stream;
    rc = batch_launch(ctx, B, s, nq, ks, mode, v);
    if (rc) return rc;
    TK_HIP(hipStreamSynchronize(s));
    PassTally tally;
    std::vector<std::pair<uint32_t, float>> reseed;
    std::vector<uint32_t> fte;
    collect(ctx, B, nq, ks, qs, nullptr, out_ids, out_scores, out_thr, reseed, fte, tally);

    const uint32_t retried = (uint32_t)reseed.size();
    if (!reseed.empty()) {
        std::vector<topk_query> sub(reseed.size());
        std::vector<uint32_t> rows(reseed.size());
        for (size_t i = 0; i < reseed.size(); i++) {
            rows[i] = reseed[i].first;
            sub[i] = qs[reseed[i].first];
        }
        const uint32_t n2 = (uint32_t)sub.size();
        rc = stage_queries(ctx, B, n2, sub.data(), ks);
        if (rc) return rc;
        BatchQuery *hq = reinterpret_cast<BatchQuery *>(B.h_in + InLayout::make(n2).q_off);
        for (uint32_t i = 0; i < n2; i++) hq[i].tau = reseed[i].second;
        rc = batch_enqueue(ctx, B, s, n2, ks, BMODE_GIVEN, v);
        if (rc) return rc;
        TK_HIP(hipStreamSynchronize(s));
        std::vector<std::pair<uint32_t, float>> again;
        collect(ctx, B, n2, ks, qs, rows.data(), out_ids, out_scores, out_thr, again, fte, tally);
        for (const auto &r : again) fte.push_back(r.first);
    }
    for (uint32_t i : fte) {
        const topk_query &Q = qs[i];
        float t = 0;
        rc = topk_fte(ctx, Q.k, (uint8_t)Q.qq, Q.weights, Q.attr, out_ids ? out_ids + (uint64_t)i * ks : nullptr,
                      out_scores ? out_scores + (uint64_t)i * ks : nullptr, &t);
        if (rc) return rc;
        if (out_thr) out_thr[i] = t;
    }
    topk_stats &S = ctx->stats;
    S.queries = nq;
    S.fallback = (uint32_t)fte.size();
    S.retried = retried;
    S.candidates = tally.candidates;
    S.partitions_scored = tally.pairs;
    S.partitions_total = (uint64_t)ctx->pta.nonempty * nq;
    S.tuple_count = tally.objs;
    S.ms_last = now_ms() - t0;
    return TOPK_OK;
}

}  // namespace topk

using namespace topk;

extern "C" {

int topk_batch(topk_ctx *ctx, uint32_t nq, const topk_query *qs, uint32_t k_stride, uint64_t *out_ids,
               float *out_scores, float *out_thresholds) {
    return batch_run(ctx, nq, qs, k_stride, out_ids, out_scores, out_thresholds, nullptr);
}

}  // extern "C"
