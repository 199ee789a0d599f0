// Exact top-k select and best-first emission over a candidate view, shared by
// the FTE, PTA and batch final kernels.
#pragma once

#include "device.hpp"
#include "internal.hpp"
#include "keys.hpp"

namespace topk {

// Candidate view over one contiguous key array.
struct FlatView {
    const uint64_t *p;
    __device__ __forceinline__ uint64_t get(uint32_t i) const { return p[i]; }
};

// Selects the exact top k of the c candidates of view v and writes them best
// first.  All threads of the block call it.  `sorted` (TOPK_MAX_K keys) may
// alias the view's storage when c <= BS*KPT (the keys are read into
// registers first).  Up to 2*BS candidates take one rank pass in LDS; more
// take a radix select (DB-bit digits in registers, or over memory) and a
// rank sort of the k selected.
constexpr uint32_t SMALL_RANK = 384;  // candidates ranked directly (no radix select)

template <int BS, int KPT, class View, int DB = 11>
__device__ void select_and_emit(const View &v, uint32_t c, uint32_t k, const FinalOut &o, uint64_t *top,
                                uint64_t *sorted, SelectScratch &ss) {
    const uint32_t tid = threadIdx.x;
    if (c <= SMALL_RANK && c <= 2u * BS) {
        // Few candidates: one rank pass (c compares per key, VALU-bound:
        // beyond a few hundred keys the radix select below is faster).
        // Keys are distinct (ids are), so the number of larger keys is each
        // key's output position.
        uint64_t mk[2];
        uint32_t r[2] = {0u, 0u};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const uint32_t i = tid + u * BS;
            mk[u] = i < c ? v.get(i) : 0ull;
        }
        __syncthreads();  // the view's storage may be reused as `sorted`
#pragma unroll
        for (int u = 0; u < 2; u++)
            if (tid + u * BS < c) sorted[tid + u * BS] = mk[u];
        __syncthreads();
        rank2_lds(sorted, c, mk[0], mk[1], r[0], r[1]);
        const uint32_t got = c < k ? c : k;
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if (tid + u * BS < c && r[u] < k) {
                o.ids[r[u]] = rkey_id(mk[u]);
                o.scores[r[u]] = rkey_score(mk[u]);
                if (r[u] == got - 1 && o.thr) *o.thr = rkey_score(mk[u]);
            }
        }
        for (uint32_t i = got + tid; i < k; i += BS) {
            o.ids[i] = ~0ull;
            o.scores[i] = -INFINITY;
        }
        if (tid == 0) {
            if (!got && o.thr) *o.thr = -INFINITY;
            if (o.status) *o.status = 0;
        }
        return;
    }
    uint32_t got;
    if (c <= (uint32_t)BS * KPT) {
        RegKeys<KPT> rk;
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            const uint32_t i = j * BS + tid;
            rk.k[j] = i < c ? v.get(i) : 0ull;
        }
        __syncthreads();  // the view's storage may be reused as `sorted` below
        const uint64_t thr = c > k ? block_kth_largest<BS, RegKeys<KPT>, DB>(rk, c, k, ss) : 1ull;
        got = block_gather_ge<BS>(rk, thr, top, ss);
    } else {
        const uint64_t thr = block_kth_largest<BS>(ViewKeys<BS, View, false>{v, c}, c, k, ss);
        got = block_gather_ge<BS>(ViewKeys<BS, View, true>{v, c}, thr, top, ss);
    }
    block_rank_sort_desc<BS>(top, got, sorted);
    for (uint32_t i = tid; i < k; i += BS) {
        if (i < got) {
            o.ids[i] = rkey_id(sorted[i]);
            o.scores[i] = rkey_score(sorted[i]);
        } else {
            o.ids[i] = ~0ull;
            o.scores[i] = -INFINITY;
        }
    }
    if (tid == 0) {
        if (o.thr) *o.thr = got ? rkey_score(sorted[got - 1]) : -INFINITY;
        if (o.status) *o.status = 0;
    }
}

}  // namespace topk
