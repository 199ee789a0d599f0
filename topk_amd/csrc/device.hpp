// Wave64 and workgroup primitives for gfx950: scans and reductions, an exact
// workgroup radix select over 64-bit ranking keys, key gathering and a rank
// sort in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "keys.hpp"

namespace topk {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t t = __shfl_xor(v, o, 64);
        v = t < v ? t : v;
    }
    return v;
}

// The last index i in [lo, hi) with key(i) <= t (key non-decreasing, key(lo)
// <= t), found by the wave: 64 probes per round narrow the range 64-fold.
template <class Key>
__device__ __forceinline__ uint32_t wave_last_le(Key key, uint32_t lo, uint32_t hi, uint32_t t) {
    const uint32_t lane = lane_id();
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t i = lo + lane * step;
        const bool le = i < hi && key(i) <= t;
        const uint64_t m = __ballot(le);  // lanes 0..j (keys are non-decreasing)
        const uint32_t j = 63u - (uint32_t)__builtin_clzll(m);
        lo = lo + j * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    return lo;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Streaming 128-bit load that bypasses cache residency (data read once).
__device__ __forceinline__ f4 ld_stream(const float *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p));
}

// LDS scratch of the workgroup selection (histogram for up to 11-bit digits).
struct SelectScratch {
    uint32_t hist[2048];
    uint32_t digit, need, bcnt, counter;
    unsigned long long kmin, kmax;
};

// How many of the n keys a[0..n) (LDS, every lane reading the same word)
// are above x and above y: eight reads in flight per step, so the loop
// runs at LDS throughput rather than latency.
__device__ __forceinline__ void rank2_lds(const uint64_t *a, uint32_t n, uint64_t x, uint64_t y, uint32_t &rx,
                                          uint32_t &ry) {
    uint32_t j = 0;
    for (; j + 8 <= n; j += 8) {
        uint64_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a[j + u];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            rx += v[u] > x ? 1u : 0u;
            ry += v[u] > y ? 1u : 0u;
        }
    }
    for (; j < n; j++) {
        const uint64_t v = a[j];
        rx += v > x ? 1u : 0u;
        ry += v > y ? 1u : 0u;
    }
}

// Key sources: the keys a thread owns, visited by for_each.  Key 0 marks an
// empty slot and is never selected.
template <int KPT>
struct RegKeys {
    uint64_t k[KPT];
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
#pragma unroll
        for (int i = 0; i < KPT; i++) f(k[i]);
    }
};

template <int BS>
struct MemKeys {
    const uint64_t *p;
    uint32_t cnt;
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
        uint32_t i = threadIdx.x;
        for (; i + 7 * BS < cnt; i += 8 * BS) {  // 8 independent loads in flight
            uint64_t v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = p[i + j * BS];
#pragma unroll
            for (int j = 0; j < 8; j++) f(v[j]);
        }
        for (; i < cnt; i += BS) f(p[i]);
    }
};

// Block-wide min / max of u64 (all threads call; result on every thread).
template <int BS>
__device__ __forceinline__ void block_minmax(uint64_t &lo, uint64_t &hi, SelectScratch &s) {
    lo = wave_min_u64(lo);
    hi = wave_max_u64(hi);
    if (BS == 64) return;
    if (threadIdx.x == 0) { s.kmin = ~0ull; s.kmax = 0; }
    __syncthreads();
    if (lane_id() == 0) {
        atomicMin(&s.kmin, (unsigned long long)lo);
        atomicMax(&s.kmax, (unsigned long long)hi);
    }
    __syncthreads();
    lo = s.kmin;
    hi = s.kmax;
    __syncthreads();
}

// Exact k-th largest non-zero key of the block's key source, where setcnt
// (>= k >= 1) is the number of non-zero keys.  Every thread of the block
// (BS = blockDim.x, a multiple of 64) must call it and gets the result.
//
// MSD radix select that never spends a pass on bits all remaining
// candidates share: each round takes the min / max of the candidate set,
// starts its 8-bit digit just below their common prefix, builds the digit
// histogram in LDS, and one wave finds the digit holding the k-th key with
// a lane prefix scan.  Rounds stop when the candidate set is one value, when
// the k-th is its maximum, or when the whole set is inside the top k (then
// the k-th is its minimum).
template <int BS, class Src, int DB = 8>
__device__ uint64_t block_kth_largest(const Src &src, uint32_t setcnt, uint32_t k, SelectScratch &s) {
    static_assert(DB >= 6 && DB <= 11, "digit width");
    constexpr uint32_t NB = 1u << DB, PER = NB / 64;
    const uint32_t tid = threadIdx.x;
    uint64_t prefix = 0, mask = 0;
    uint32_t need = k;
    for (;;) {
        uint64_t lo = ~0ull, hi = 0;
        src.for_each([&](uint64_t key) {
            if (key && (key & mask) == prefix) {
                lo = key < lo ? key : lo;
                hi = key > hi ? key : hi;
            }
        });
        block_minmax<BS>(lo, hi, s);
        if (need == setcnt) return lo;
        if (need == 1 || lo == hi) return hi;
        const int db = 63 - __clzll(lo ^ hi);          // highest differing bit
        const uint64_t above = db == 63 ? 0ull : (~0ull << (db + 1));
        prefix = hi & above;
        mask = above;
        const int width = db + 1 < DB ? db + 1 : DB;
        const int shift = db + 1 - width;
        const uint32_t dmask = (1u << width) - 1u;
        for (uint32_t i = tid; i < NB; i += BS) s.hist[i] = 0;
        __syncthreads();
        src.for_each([&](uint64_t key) {
            if (key && (key & mask) == prefix) atomicAdd(&s.hist[(uint32_t)(key >> shift) & dmask], 1u);
        });
        __syncthreads();
        if (tid < 64) {
            // lane l owns bins NB-1-PER*l .. NB-PER*(l+1), highest first
            const uint32_t top = NB - 1 - PER * tid;
            uint32_t tot = 0;
            for (uint32_t j = 0; j < PER; j++) tot += s.hist[top - j];
            const uint32_t incl = wave_incl_scan(tot);
            const uint32_t excl = incl - tot;
            if (excl < need && need <= incl) {
                uint32_t r = need - excl, acc = 0;
                for (uint32_t j = 0; j < PER; j++) {
                    const uint32_t cj = s.hist[top - j];
                    if (r <= acc + cj) {
                        s.digit = top - j;
                        s.need = r - acc;
                        s.bcnt = cj;
                        break;
                    }
                    acc += cj;
                }
            }
        }
        __syncthreads();
        const uint32_t digit = s.digit;
        need = s.need;
        setcnt = s.bcnt;
        __syncthreads();
        prefix |= (uint64_t)digit << shift;
        mask |= (uint64_t)dmask << shift;
    }
}

// Writes every non-zero key >= thr of the source to out[] (order
// unspecified) with a wave-aggregated LDS counter; returns the count.  All
// threads of the block must call.
template <int BS, class Src, class Out>
__device__ uint32_t block_gather_ge(const Src &src, uint64_t thr, Out out, SelectScratch &s) {
    if (threadIdx.x == 0) s.counter = 0;
    __syncthreads();
    src.for_each([&](uint64_t key) {
        const bool p = key && key >= thr;
        const uint64_t bal = __ballot(p);
        if (bal) {
            const uint32_t rank = __popcll(bal & ((1ull << lane_id()) - 1ull));
            uint32_t base = 0;
            if (lane_id() == 0) base = atomicAdd(&s.counter, (uint32_t)__popcll(bal));
            base = __shfl(base, 0, 64);
            if (p) out[base + rank] = key;
        }
    });
    __syncthreads();
    const uint32_t r = s.counter;
    __syncthreads();
    return r;
}

// Memory key source for block_gather_ge: every lane of a wave must take
// part in the ballot, so the trip count is uniform across the block.
template <int BS>
struct MemKeysUniform {
    const uint64_t *p;
    uint32_t cnt;
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
        for (uint32_t i0 = 0; i0 < cnt; i0 += BS) {
            const uint32_t i = i0 + threadIdx.x;
            f(i < cnt ? p[i] : 0ull);
        }
    }
};

// Key source over a flat index space [0, cnt) resolved by a view's get(p);
// Uniform=true keeps the trip count uniform across the block (for ballots).
template <int BS, class View, bool Uniform>
struct ViewKeys {
    View v;
    uint32_t cnt;
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
        if constexpr (Uniform) {
            for (uint32_t i0 = 0; i0 < cnt; i0 += BS) {
                const uint32_t i = i0 + threadIdx.x;
                f(i < cnt ? v.get(i) : 0ull);
            }
        } else {
            for (uint32_t i = threadIdx.x; i < cnt; i += BS) f(v.get(i));
        }
    }
};

// Block-wide exclusive prefix sum of one value per thread; returns the
// thread's exclusive prefix and writes the block total to *total.
template <int BS>
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *wave_sums, uint32_t *total) {
    constexpr int NW = BS / 64;
    const uint32_t w = threadIdx.x / 64;
    const uint32_t incl = wave_incl_scan(v);
    if (lane_id() == 63) wave_sums[w] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t ws = threadIdx.x < NW ? wave_sums[threadIdx.x] : 0u;
        const uint32_t wi = wave_incl_scan(ws);
        if (threadIdx.x < NW) wave_sums[threadIdx.x] = wi - ws;
        if (threadIdx.x == NW - 1) *total = wi;
    }
    __syncthreads();
    const uint32_t r = wave_sums[w] + incl - v;
    __syncthreads();
    return r;
}

// The kth (1 <= kth <= 8) largest of the keys the wave's lanes hold (NPL
// per lane; distinct except zeros), by kth rounds of a wave max.  Every lane
// gets the result; keys below it stay in v, the chosen ones are zeroed.
template <int NPL>
__device__ __forceinline__ uint64_t wave_kth_largest(uint64_t (&v)[NPL], uint32_t kth) {
    uint64_t mx = 0;
    for (uint32_t r = 0; r < kth; r++) {
        uint64_t lm = 0;
#pragma unroll
        for (int j = 0; j < NPL; j++) lm = v[j] > lm ? v[j] : lm;
        mx = wave_max_u64(lm);
#pragma unroll
        for (int j = 0; j < NPL; j++) v[j] = v[j] == mx ? 0ull : v[j];
    }
    return mx;
}

// Sorts the got distinct keys in[0..got) descending into out[] by rank
// counting (one pass, no barrier inside): out[#{j: in[j] > in[i]}] = in[i].
template <int BS>
__device__ void block_rank_sort_desc(const uint64_t *in, uint32_t got, uint64_t *out) {
    for (uint32_t i = threadIdx.x; i < got; i += BS) {
        const uint64_t x = in[i];
        uint32_t r = 0, unused = 0;
        rank2_lds(in, got, x, x, r, unused);
        out[r] = x;
    }
    __syncthreads();
}

// Sorts a[0..n2) descending in LDS (n2 a power of two); all threads call.
template <int BS>
__device__ void bitonic_sort_desc(uint64_t *a, uint32_t n2) {
    for (uint32_t size = 2; size <= n2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n2 / 2; t += BS) {
                uint32_t lo = 2 * stride * (t / stride) + (t & (stride - 1));
                uint32_t hi = lo + stride;
                bool desc = (lo & size) == 0;
                uint64_t x = a[lo], y = a[hi];
                if ((x < y) == desc) { a[lo] = y; a[hi] = x; }
            }
            __syncthreads();
        }
    }
}

__host__ __device__ __forceinline__ uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace topk
