// Wave64 and workgroup primitives for gfx950: scans, an exact LDS radix
// select over 64-bit ranking keys, and a bitonic sort in LDS.
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

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Streaming 128-bit load that bypasses cache residency (data read once).
__device__ __forceinline__ f4 ld_stream(const float *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p));
}

// LDS scratch of the block-wide selection.
struct SelectScratch {
    uint32_t hist[256];
    uint32_t digit, need, bcnt, flag;
    unsigned long long kmin;
};

// Exact k-th largest of keys[0..cnt) (1 <= k <= cnt), computed by every
// thread of the block (all threads must call; BS = blockDim.x, a multiple of
// 64).  MSD radix select over 8-bit digits with LDS histograms; the digit
// holding the k-th key is found by one wave with a lane prefix scan.  When
// the bucket holding the k-th key lies entirely inside the top k, the
// remaining digits are skipped and the bucket minimum is returned.  Keys may
// repeat; the returned value is still the k-th largest.
template <int BS, typename KeyPtr>
__device__ uint64_t block_kth_largest(KeyPtr keys, uint32_t cnt, uint32_t k, SelectScratch &s) {
    const uint32_t tid = threadIdx.x;
    uint64_t prefix = 0, mask = 0;
    uint32_t need = k;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (uint32_t i = tid; i < 256; i += BS) s.hist[i] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < cnt; i += BS) {
            uint64_t key = keys[i];
            if ((key & mask) == prefix) atomicAdd(&s.hist[(uint32_t)(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            // lane l owns bins 255-4l .. 252-4l, highest first
            uint32_t c0 = s.hist[255 - 4 * tid], c1 = s.hist[254 - 4 * tid];
            uint32_t c2 = s.hist[253 - 4 * tid], c3 = s.hist[252 - 4 * tid];
            uint32_t tot = c0 + c1 + c2 + c3;
            uint32_t incl = wave_incl_scan(tot);
            uint32_t excl = incl - tot;
            if (excl < need && need <= incl) {
                uint32_t r = need - excl, acc = 0;
                uint32_t cs[4] = {c0, c1, c2, c3};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (r <= acc + cs[j]) {
                        s.digit = 255 - 4 * tid - j;
                        s.need = r - acc;
                        s.bcnt = cs[j];
                        break;
                    }
                    acc += cs[j];
                }
            }
        }
        __syncthreads();
        const uint32_t digit = s.digit, bcnt = s.bcnt;
        need = s.need;
        prefix |= (uint64_t)digit << shift;
        mask |= (uint64_t)0xFF << shift;
        if (bcnt == need) {
            if (tid == 0) s.kmin = ~0ull;
            __syncthreads();
            uint64_t lm = ~0ull;
            for (uint32_t i = tid; i < cnt; i += BS) {
                uint64_t key = keys[i];
                if ((key & mask) == prefix && key < lm) lm = key;
            }
            lm = wave_min_u64(lm);
            if (lane_id() == 0) atomicMin(&s.kmin, (unsigned long long)lm);
            __syncthreads();
            uint64_t r = s.kmin;
            __syncthreads();
            return r;
        }
    }
    __syncthreads();
    return prefix;
}

// Appends every key >= thr from keys[0..cnt) to out[] (order unspecified)
// using a wave-aggregated LDS counter; returns the number appended.  All
// threads of the block must call.
template <int BS, typename KeyPtr, typename OutPtr>
__device__ uint32_t block_gather_ge(KeyPtr keys, uint32_t cnt, uint64_t thr, OutPtr out,
                                    uint32_t *lds_counter) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) *lds_counter = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < cnt; i0 += BS) {
        uint32_t i = i0 + tid;
        uint64_t key = i < cnt ? keys[i] : 0;
        bool p = (i < cnt) && key >= thr;
        uint64_t bal = __ballot(p);
        if (bal) {
            uint32_t nw = __popcll(bal);
            uint32_t rank = __popcll(bal & ((1ull << lane_id()) - 1ull));
            uint32_t base = 0;
            if (lane_id() == 0) base = atomicAdd(lds_counter, nw);
            base = __shfl(base, 0, 64);
            if (p) out[base + rank] = key;
        }
    }
    __syncthreads();
    uint32_t r = *lds_counter;
    __syncthreads();
    return r;
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

__device__ __forceinline__ uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace topk
