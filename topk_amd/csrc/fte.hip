// FTE (full-table evaluation) top-k on gfx950.
//
// Reference: cpu/TPAc.h:118-358 (scoring order, heap semantics) and the
// empty GPU stub gpu/GFA.h:47-50.  The pipeline is three launches on the
// context's stream, with no host round trip:
//
//  1. prefix pass  (fte_prefix_kernel): the first P = G1*4096 objects are
//     scored, each 4096-object block keeps its exact top-k (LDS radix
//     select), and the last block to finish merges the G1*k block winners
//     into the prefix's exact top-k.  Its k-th key tau is a safe lower bound
//     of the global k-th key (k objects reach it).
//  2. stream pass  (fte_stream_kernel): objects [P, n) are scored with
//     128-bit column loads, scores stay in registers, and only keys > tau
//     are appended (wave ballot + one atomic per wave) to a candidate list.
//     No barrier and no LDS in the loop: a pure HBM stream.
//  3. final select (fte_final_kernel): one workgroup selects the exact top-k
//     of the candidates and bitonic-sorts them best first.
//
// Exactness: every object outside the candidate list has key <= tau <=
// global k-th key and is a different object from the k-th, so it is not in
// the top k.  If the list overflows its capacity the final kernel flags
// STATUS_OVERFLOW and the host re-runs with a buffer of n entries.
#include "device.hpp"
#include "internal.hpp"

namespace topk {

constexpr int PREFIX_BS = 256;
constexpr uint32_t PREFIX_TILE = 4096;   // objects per prefix block
constexpr int STREAM_BS = 256;
constexpr int STREAM_UNR = 2;            // float4 vectors per thread per step
constexpr int FINAL_BS = 1024;

// Score of the 4 objects of float4 vector v: acc = fl(acc + fl(x*w)) in
// attribute order, accumulator starting at +0 (cpu/TPAc.h:129-142).
template <int M>
__device__ __forceinline__ f4 score_vec(const QCols &q, uint64_t v) {
    const uint32_t m = M ? M : q.m;
    f4 x[M ? M : 1];
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (M) {
#pragma unroll
        for (int j = 0; j < (M ? M : 1); j++) x[j] = ld_stream(q.col[j] + v * 4);
#pragma unroll
        for (int j = 0; j < (M ? M : 1); j++) {
            f4 p = x[j] * q.w[j];
            acc = acc + p;
        }
    } else {
        for (uint32_t j = 0; j < m; j++) {
            f4 xv = ld_stream(q.col[j] + v * 4);
            f4 p = xv * q.w[j];
            acc = acc + p;
        }
    }
    return acc;
}

// ---------------------------------------------------------------------------
// 1. prefix pass
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(PREFIX_BS) void fte_prefix_kernel(QCols q, uint64_t n_eff, uint32_t k,
                                                               uint64_t *__restrict__ blk,
                                                               uint64_t *__restrict__ cand,
                                                               FteState *__restrict__ st) {
    __shared__ uint64_t keys[PREFIX_TILE];
    __shared__ SelectScratch ss;
    __shared__ uint32_t counter;
    __shared__ uint32_t is_last;
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * PREFIX_TILE;

#pragma unroll
    for (int j = 0; j < (int)(PREFIX_TILE / (4 * PREFIX_BS)); j++) {
        const uint32_t lv = j * PREFIX_BS + tid;       // local float4 index
        const uint64_t v = base / 4 + lv;
        f4 s = score_vec<M>(q, v);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            uint64_t id = v * 4 + e;
            keys[lv * 4 + e] = id < n_eff ? rkey(s[e], (uint32_t)id) : 0ull;
        }
    }
    __syncthreads();

    const uint64_t valid64 = n_eff > base ? (n_eff - base < PREFIX_TILE ? n_eff - base : PREFIX_TILE) : 0;
    const uint32_t valid = (uint32_t)valid64;
    uint64_t thr = 1;
    if (valid > k) thr = block_kth_largest<PREFIX_BS>(keys, PREFIX_TILE, k, ss);
    uint64_t *out = blk + (uint64_t)blockIdx.x * k;
    uint32_t got = block_gather_ge<PREFIX_BS>(keys, PREFIX_TILE, thr, out, &counter);
    for (uint32_t i = got + tid; i < k; i += PREFIX_BS) out[i] = 0;

    // last block to arrive merges the block winners
    __threadfence();
    __syncthreads();
    if (tid == 0) is_last = (atomicAdd(&st->arrive, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!is_last) return;
    __threadfence();

    const uint32_t g = gridDim.x;
    const uint64_t full = n_eff / PREFIX_TILE, rem = n_eff % PREFIX_TILE;
    const uint64_t total = full * (k < PREFIX_TILE ? k : PREFIX_TILE) + (rem < k ? rem : k);
    uint64_t tau = 1;
    if (total > k) tau = block_kth_largest<PREFIX_BS>(blk, g * k, k, ss);
    uint32_t c = block_gather_ge<PREFIX_BS>(blk, g * k, tau, cand, &counter);
    if (tid == 0) {
        st->tau = tau;
        st->count = c;
        st->arrive = 0;
    }
}

// ---------------------------------------------------------------------------
// 2. stream pass
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(STREAM_BS) void fte_stream_kernel(QCols q, uint64_t n, uint64_t v_begin,
                                                               uint64_t v_end,
                                                               FteState *__restrict__ st,
                                                               uint64_t *__restrict__ cand,
                                                               uint32_t cap) {
    const uint64_t tau = st->tau;
    const uint32_t tau_hi = (uint32_t)(tau >> 32);
    const uint64_t stride = (uint64_t)gridDim.x * STREAM_BS;
    const uint32_t lane = lane_id();
    // The trip count depends on blockIdx only, so every lane of a wave stays
    // active through the ballot / scan below; lanes past v_end are masked.
    for (uint64_t vb = v_begin + (uint64_t)blockIdx.x * STREAM_BS; vb < v_end; vb += stride * STREAM_UNR) {
        const uint64_t v0 = vb + threadIdx.x;
        f4 s[STREAM_UNR];
#pragma unroll
        for (int u = 0; u < STREAM_UNR; u++) {
            const uint64_t vv = v0 + u * stride;
            if (vv < v_end) s[u] = score_vec<M>(q, vv);
            else s[u] = f4{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
        }
        // filter: key > tau (score key first; ids only on ties / passes)
        uint32_t npass = 0;
        uint64_t kk[STREAM_UNR * 4];
#pragma unroll
        for (int u = 0; u < STREAM_UNR; u++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint64_t id = (v0 + u * stride) * 4 + e;
                const uint32_t ok = okey(s[u][e]);
                const uint64_t key = rkey_from_okey(ok, (uint32_t)id);
                const bool p = ok >= tau_hi && key > tau && id < n;
                kk[u * 4 + e] = p ? key : 0ull;
                npass += p ? 1u : 0u;
            }
        }
        if (__ballot(npass != 0)) {
            const uint32_t incl = wave_incl_scan(npass);
            uint32_t base = 0;
            if (lane == 63) base = atomicAdd(&st->count, incl);
            base = __shfl(base, 63, 64);
            uint32_t pos = base + incl - npass;
#pragma unroll
            for (int j = 0; j < STREAM_UNR * 4; j++) {
                if (kk[j]) {
                    if (pos < cap) cand[pos] = kk[j];
                    pos++;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 3. final select + sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FINAL_BS) void fte_final_kernel(FteState *__restrict__ st,
                                                             const uint64_t *__restrict__ cand,
                                                             uint32_t cap, uint32_t k,
                                                             uint64_t *__restrict__ out_ids,
                                                             float *__restrict__ out_scores,
                                                             float *__restrict__ out_thr,
                                                             uint32_t *__restrict__ out_status) {
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ SelectScratch ss;
    __shared__ uint32_t counter;
    const uint32_t tid = threadIdx.x;
    const uint32_t c = st->count;
    if (c > cap) {
        if (tid == 0) {
            st->status |= STATUS_OVERFLOW;
            if (out_status) *out_status = STATUS_OVERFLOW;
        }
        return;
    }
    uint64_t thr = 1;
    if (c > k) thr = block_kth_largest<FINAL_BS>(cand, c, k, ss);
    const uint32_t got = block_gather_ge<FINAL_BS>(cand, c, thr, top, &counter);
    const uint32_t n2 = next_pow2(k > 1 ? k : 2);
    for (uint32_t i = got + tid; i < n2; i += FINAL_BS) top[i] = 0;
    __syncthreads();
    bitonic_sort_desc<FINAL_BS>(top, n2);
    for (uint32_t i = tid; i < k; i += FINAL_BS) {
        if (i < got) {
            out_ids[i] = rkey_id(top[i]);
            out_scores[i] = rkey_score(top[i]);
        } else {
            out_ids[i] = ~0ull;
            out_scores[i] = -INFINITY;
        }
    }
    if (tid == 0) {
        if (out_thr) *out_thr = got ? rkey_score(top[got - 1]) : -INFINITY;
        if (out_status) *out_status = 0;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Launches KERN<M> for the query width m (M = 0: runtime-width fallback).
#define TK_LAUNCH_M(m, KERN, grid, block, shmem, stream, ...)                                \
    switch (m) {                                                                              \
        case 1: hipLaunchKernelGGL(KERN<1>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 2: hipLaunchKernelGGL(KERN<2>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 3: hipLaunchKernelGGL(KERN<3>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 4: hipLaunchKernelGGL(KERN<4>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 5: hipLaunchKernelGGL(KERN<5>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 6: hipLaunchKernelGGL(KERN<6>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 7: hipLaunchKernelGGL(KERN<7>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL(KERN<8>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 9: hipLaunchKernelGGL(KERN<9>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        case 10: hipLaunchKernelGGL(KERN<10>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 11: hipLaunchKernelGGL(KERN<11>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 12: hipLaunchKernelGGL(KERN<12>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 13: hipLaunchKernelGGL(KERN<13>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 14: hipLaunchKernelGGL(KERN<14>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 15: hipLaunchKernelGGL(KERN<15>, grid, block, shmem, stream, __VA_ARGS__); break; \
        case 16: hipLaunchKernelGGL(KERN<16>, grid, block, shmem, stream, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERN<0>, grid, block, shmem, stream, __VA_ARGS__); break;  \
    }

uint32_t fte_prefix_blocks(uint64_t n, uint32_t k) {
    uint64_t need = (n + PREFIX_TILE - 1) / PREFIX_TILE;
    uint32_t g = 32768u / (k ? k : 1);
    if (g > 256) g = 256;
    if (g < 8) g = 8;
    return (uint32_t)(need < g ? need : g);
}

uint32_t stream_grid(uint64_t nvec) {
    uint64_t blocks = (nvec + (uint64_t)STREAM_BS * STREAM_UNR - 1) / ((uint64_t)STREAM_BS * STREAM_UNR);
    const uint64_t cap = 256 * 16;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return (uint32_t)blocks;
}

int fte_enqueue(topk_ctx *ctx, uint32_t k, const QCols &q, uint64_t *d_ids, float *d_scores,
                float *d_thr, uint32_t *d_status, bool big) {
    const uint64_t n = ctx->n;
    const uint32_t g1 = fte_prefix_blocks(n, k);
    const uint64_t p = (uint64_t)g1 * PREFIX_TILE;
    const uint64_t n_eff = n < p ? n : p;
    uint64_t *cand = big ? ctx->cand_big : ctx->cand;
    const uint32_t cap = big ? (uint32_t)(ctx->cand_big_cap > 0xFFFFFFFFull ? 0xFFFFFFFFu : ctx->cand_big_cap)
                             : ctx->cand_cap;
    hipStream_t s = ctx->stream;

    TK_LAUNCH_M(q.m, fte_prefix_kernel, dim3(g1), dim3(PREFIX_BS), 0, s, q, n_eff, k, ctx->blk, cand,
                ctx->st);
    if (n > p) {
        const uint64_t v_begin = p / 4, v_end = (n + 3) / 4;
        const uint32_t g2 = stream_grid(v_end - v_begin);
        hipEvent_t ev0 = nullptr;
        timing_begin(ctx, "fte_stream", ev0);
        TK_LAUNCH_M(q.m, fte_stream_kernel, dim3(g2), dim3(STREAM_BS), 0, s, q, n, v_begin, v_end, ctx->st,
                    cand, cap);
        timing_end(ctx, "fte_stream", ev0);
    }
    hipLaunchKernelGGL(fte_final_kernel, dim3(1), dim3(FINAL_BS), 0, s, ctx->st, cand, cap, k, d_ids, d_scores,
                       d_thr, d_status);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

}  // namespace topk
