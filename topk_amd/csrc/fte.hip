// FTE (full-table evaluation) top-k on gfx950.
//
// Reference: cpu/TPAc.h:118-358 (scoring order, heap semantics) and the
// empty GPU stub gpu/GFA.h:47-50.  A query is three launches on the
// context's stream with no host round trip:
//
//  1. seed pass    (fte_seed_kernel): a few 4096-object chunks are scored;
//     each block keeps its top ks keys (register-resident radix select) and
//     the last block to finish selects the ks-th largest of the block
//     winners: tau.
//       - sample mode: 16 chunks spread over the id range, small ks; tau is
//         a guess whose safety is checked at the end (>= k candidates).
//       - exact mode:  contiguous chunks [0, P), ks = k; the prefix's k-th
//         key is a proven lower bound of the global k-th key, and the
//         prefix's top k are written to the candidate list.
//  2. stream pass  (fte_stream_kernel): each block streams whole tiles of
//     contiguous column segments, scores stay in registers, and keys >= tau
//     go to the block's own candidate slots (wave ballot + one LDS atomic
//     per wave; no global atomic).  No barrier in the loop: a pure HBM
//     stream.
//  3. final select (fte_final_kernel): one workgroup scans the per-block
//     counts, checks the candidate total, selects the exact top k and
//     rank-sorts them.
//
// Exactness: with c >= k candidates of key >= tau, every other object has a
// key < tau and is beaten by c >= k objects, so the top k are among the
// candidates.  Too few candidates (a sample guess that was too high) raise
// STATUS_RETRY, too many for the buffer STATUS_OVERFLOW; the host then runs
// the exact mode, with an n-entry buffer if needed.
#include <string.h>

#include <cmath>
#include <mutex>

#include "device.hpp"
#include "emit.hpp"
#include "internal.hpp"

namespace topk {

constexpr int SEED_BS = 256;
constexpr uint32_t SEED_TILE = 4096;     // exact mode: objects per seed block (16 per thread)
constexpr uint32_t SAMPLE_TILE = 1024;   // sample mode: objects per seed block (4 per thread)
constexpr uint32_t SAMPLE_CHUNKS = 128;  // sample mode: 128 x 1024 objects spread over the ids
constexpr int STREAM_BS = 256;
constexpr int FINAL_BS = 512;            // regular final select
constexpr int FINAL_KPT = 16;            // candidate keys per final-select thread (8192 in registers)
constexpr uint32_t SEED_WAVE_MAX = 512;  // seed winners a stream wave can reduce to tau itself
constexpr int FLAT_BS = 1024;            // fallback final select over a large flat list

// Score of the 4 objects of float4 vector v: acc = fl(acc + fl(x*w)) in
// attribute order, accumulator starting at +0 (cpu/TPAc.h:129-142).
template <int M>
__device__ __forceinline__ f4 score_vec(const QCols &q, uint64_t v) {
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (M > 0) {
        f4 x[M];
#pragma unroll
        for (int j = 0; j < M; j++) x[j] = ld_stream(q.col[j] + v * 4);
#pragma unroll
        for (int j = 0; j < M; j++) {
            f4 p = x[j] * q.w[j];
            acc = acc + p;
        }
    } else {
        for (uint32_t j = 0; j < q.m; j++) {
            f4 xv = ld_stream(q.col[j] + v * 4);
            f4 p = xv * q.w[j];
            acc = acc + p;
            if (q.div != 0.0f) acc = acc / q.div;
        }
    }
    return acc;
}

// ---------------------------------------------------------------------------
// 1. seed pass
// ---------------------------------------------------------------------------
// Block b scores objects [b*chunk_stride, +256*OPT) of the first n_eff objects.
// total_valid = number of non-zero keys the blocks deposit in blk.
template <int M, int OPT, bool MERGE>
__global__ __launch_bounds__(SEED_BS) void fte_seed_kernel(QCols q, uint64_t n_eff, uint64_t chunk_stride,
                                                           uint32_t ks, uint32_t total_valid,
                                                           uint64_t *__restrict__ blk,
                                                           uint64_t *__restrict__ pre,
                                                           FteState *__restrict__ st,
                                                           const uint32_t *__restrict__ gate) {
    __shared__ SelectScratch ss;
    __shared__ uint32_t is_last;
    if (gate && *gate == 0) return;  // the pass this one backs up settled the query
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * chunk_stride;

    constexpr uint32_t TILE = SEED_BS * OPT;
    RegKeys<OPT> rk;
#pragma unroll
    for (int j = 0; j < OPT / 4; j++) {
        const uint64_t v = base / 4 + j * SEED_BS + tid;
        const f4 s = score_vec<M>(q, v);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const uint64_t id = v * 4 + e;
            rk.k[j * 4 + e] = id < n_eff ? rkey(s[e], (uint32_t)id) : 0ull;
        }
    }
    const uint64_t left = n_eff > base ? n_eff - base : 0;
    const uint32_t valid = (uint32_t)(left < TILE ? left : TILE);
    uint64_t thr = 1;
    if (valid > ks) thr = block_kth_largest<SEED_BS>(rk, valid, ks, ss);
    uint64_t *out = blk + (uint64_t)blockIdx.x * ks;
    const uint32_t got = block_gather_ge<SEED_BS>(rk, thr, out, ss);
    for (uint32_t i = got + tid; i < ks; i += SEED_BS) out[i] = 0;
    if constexpr (!MERGE) return;  // sample mode: each stream wave derives tau from blk itself

    // the last block to arrive merges the block winners
    __threadfence();
    __syncthreads();
    if (tid == 0) is_last = (atomicAdd(&st->arrive, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!is_last) return;
    __threadfence();

    const uint32_t nk = gridDim.x * ks;
    uint64_t tau = 1;
    if (total_valid > ks) tau = block_kth_largest<SEED_BS>(MemKeys<SEED_BS>{blk, nk}, total_valid, ks, ss);
    uint32_t c = 0;
    if (pre) c = block_gather_ge<SEED_BS>(MemKeysUniform<SEED_BS>{blk, nk}, tau, pre, ss);
    if (tid == 0) {
        st->tau = tau;
        st->pre_count = c;
        st->arrive = 0;
    }
}

// ---------------------------------------------------------------------------
// final select + sort
// ---------------------------------------------------------------------------
// Flat index -> key over pre[0..npre) followed by the blocks' slot regions;
// off[] (LDS) holds the exclusive prefix of the per-block counts.
struct SlotView {
    const uint64_t *pre;
    uint32_t npre;
    const uint64_t *keys;
    uint32_t slot;
    const uint32_t *off;   // nblk + 2 entries; region nblk is the spill region
    uint32_t nblk;
    const uint64_t *spill;
    __device__ __forceinline__ uint64_t get(uint32_t p) const {
        if (p < npre) return pre[p];
        const uint32_t q = p - npre;
        if (q >= off[nblk]) return spill[q - off[nblk]];
        uint32_t lo = 0, hi = nblk;  // largest b with off[b] <= q
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= q) lo = mid;
            else hi = mid;
        }
        return keys[(uint64_t)lo * slot + (q - off[lo])];
    }
};


// The ks-th largest of the nseed (<= 512) seed winners, computed by one wave.
__device__ __forceinline__ uint64_t wave_seed_tau(const uint64_t *__restrict__ seed, uint32_t nseed,
                                                  uint32_t ks) {
    uint64_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t i = j * 64 + lane_id();
        v[j] = i < nseed ? seed[i] : 0ull;
    }
    return wave_kth_largest<8>(v, ks);
}

// Sample mode seed with few winners per block (ks <= 4): block b scores the
// 1024 objects at b*chunk_stride, every wave keeps its top ks by wave-max
// rounds, and wave 0 reduces the 4*ks wave winners to the block's top ks.
// No radix select, one barrier.
template <int M>
__global__ __launch_bounds__(SEED_BS) void fte_sample_kernel(QCols q, uint64_t n, uint64_t chunk_stride,
                                                             uint32_t ks, uint64_t *__restrict__ blk) {
    __shared__ uint64_t wtop[SEED_BS / 64][4];
    const uint32_t tid = threadIdx.x, w = tid / 64;
    const uint64_t v = (uint64_t)blockIdx.x * chunk_stride / 4 + tid;
    const f4 s = score_vec<M>(q, v);
    uint64_t keys[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint64_t id = v * 4 + e;
        keys[e] = id < n ? rkey(s[e], (uint32_t)id) : 0ull;
    }
    for (uint32_t r = 0; r < ks; r++) {
        uint64_t lm = 0;
#pragma unroll
        for (int e = 0; e < 4; e++) lm = keys[e] > lm ? keys[e] : lm;
        const uint64_t mx = wave_max_u64(lm);
#pragma unroll
        for (int e = 0; e < 4; e++) keys[e] = keys[e] == mx ? 0ull : keys[e];
        if (lane_id() == 0) wtop[w][r] = mx;
    }
    __syncthreads();
    if (w == 0) {
        uint64_t c[1];
        const uint32_t i = lane_id();
        c[0] = i < (SEED_BS / 64) * ks ? wtop[i / ks][i % ks] : 0ull;
        for (uint32_t r = 0; r < ks; r++) {
            const uint64_t mx = wave_max_u64(c[0]);
            c[0] = c[0] == mx ? 0ull : c[0];
            if (i == 0) blk[(uint64_t)blockIdx.x * ks + r] = mx;
        }
    }
}

// ---------------------------------------------------------------------------
// 2. stream pass
// ---------------------------------------------------------------------------
// Each block streams whole tiles of TILE = 1024*VPT objects: for every
// queried column it reads one contiguous TILE*4-byte segment (VPT float4 per
// lane).  Block-contiguous segments keep HBM near 88% of peak where a
// thread-level grid stride over the columns stays near 72%.  The column
// store is padded by one tile, so tile loads need no bounds check.
// tau comes from st->tau (published by a merging seed pass) or, when nseed
// > 0, from the ks-th largest of the nseed seed winners in seed[].
template <int M, int VPT>
__global__ __launch_bounds__(STREAM_BS) void fte_stream_kernel(QCols q, uint64_t n, uint64_t obj_begin,
                                                               uint64_t ntiles, FteState *__restrict__ st,
                                                               CandSet cs, const uint64_t *__restrict__ seed,
                                                               uint32_t nseed, uint32_t ks,
                                                               const uint32_t *__restrict__ gate) {
    constexpr uint32_t TILE = STREAM_BS * 4 * VPT;
    __shared__ uint32_t lcount;
    if (gate && *gate == 0) return;
    if (threadIdx.x == 0) lcount = 0;
    __syncthreads();
    const uint64_t tau = nseed ? wave_seed_tau(seed, nseed, ks) : st->tau;
    const uint32_t tau_hi = (uint32_t)(tau >> 32);
    const uint32_t lane = lane_id();
    uint64_t *my = cs.keys + (uint64_t)blockIdx.x * cs.slot;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t v0 = (obj_begin + t * TILE) / 4 + threadIdx.x;
        f4 s[VPT];
        if constexpr (M > 0) {
            f4 x[M][VPT];
#pragma unroll
            for (int j = 0; j < M; j++)
#pragma unroll
                for (int u = 0; u < VPT; u++) x[j][u] = ld_stream(q.col[j] + (v0 + u * STREAM_BS) * 4);
#pragma unroll
            for (int u = 0; u < VPT; u++) {
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int j = 0; j < M; j++) {
                    f4 p = x[j][u] * q.w[j];
                    acc = acc + p;
                }
                s[u] = acc;
            }
        } else {
#pragma unroll
            for (int u = 0; u < VPT; u++) s[u] = score_vec<0>(q, v0 + u * STREAM_BS);
        }
        // filter: key >= tau (score key first; the id only matters on ties)
        uint32_t npass = 0;
        uint64_t kk[VPT * 4];
#pragma unroll
        for (int u = 0; u < VPT; u++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint64_t id = (v0 + u * STREAM_BS) * 4 + e;
                const uint32_t ok = okey(s[u][e]);
                const uint64_t key = rkey_from_okey(ok, (uint32_t)id);
                const bool p = ok >= tau_hi && key >= tau && id < n;
                kk[u * 4 + e] = p ? key : 0ull;
                npass += p ? 1u : 0u;
            }
        }
        if (__ballot(npass != 0)) {
            const uint32_t incl = wave_incl_scan(npass);
            uint32_t base = 0;
            if (lane == 63) base = atomicAdd(&lcount, incl);
            base = __shfl(base, 63, 64);
            const uint32_t wtotal = __shfl(incl, 63, 64);
            // positions >= slot spill to the shared region (one global atomic per wave)
            const uint32_t first_spill = base > cs.slot ? base : cs.slot;
            const uint32_t nspill = base + wtotal > first_spill ? base + wtotal - first_spill : 0u;
            uint32_t sbase = 0;
            if (nspill) {
                if (lane == 63) sbase = atomicAdd(&st->spill, nspill);
                sbase = __shfl(sbase, 63, 64);
            }
            uint32_t pos = base + incl - npass;
#pragma unroll
            for (int j = 0; j < VPT * 4; j++) {
                if (kk[j]) {
                    if (pos < cs.slot) {
                        my[pos] = kk[j];
                    } else {
                        const uint32_t sp = sbase + (pos - first_spill);
                        if (sp < cs.spill_cap) cs.spill[sp] = kk[j];
                    }
                    pos++;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) cs.counts[blockIdx.x] = lcount < cs.slot ? lcount : cs.slot;
}

// 3. final select: one workgroup after the stream pass.  (A last-block merge
// inside the stream pass would need an agent-scope release fence in every
// one of its ~2000 blocks; each is an L2 write-back on gfx950 and together
// they cost more than a launch.)
__global__ __launch_bounds__(FINAL_BS) void fte_final_kernel(FteState *__restrict__ st, CandSet cs, uint32_t k,
                                                             uint32_t kreq, FinalOut fo,
                                                             const uint32_t *__restrict__ gate) {
    __shared__ uint32_t off[MAX_STREAM_GRID + 2];
    if (gate && *gate == 0) return;
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ uint64_t sorted[TOPK_MAX_K];
    __shared__ uint32_t wsum[FINAL_BS / 64];
    __shared__ uint32_t total;
    __shared__ SelectScratch ss;
    const uint32_t tid = threadIdx.x;
    // per-block counts -> LDS (coalesced), then exclusive offsets in place
    for (uint32_t b = tid; b < cs.nblk; b += FINAL_BS) off[b] = cs.counts[b];
    __syncthreads();
    const uint32_t per = (cs.nblk + FINAL_BS - 1) / FINAL_BS;
    const uint32_t b0 = tid * per < cs.nblk ? tid * per : cs.nblk;
    const uint32_t b1 = b0 + per < cs.nblk ? b0 + per : cs.nblk;
    uint32_t mine = 0;
    for (uint32_t b = b0; b < b1; b++) mine += off[b];
    uint32_t run = block_excl_scan<FINAL_BS>(mine, wsum, &total);
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t c = off[b];
        off[b] = run;
        run += c;
    }
    const uint32_t spilled = st->spill;
    const uint32_t nspill = spilled < cs.spill_cap ? spilled : cs.spill_cap;
    if (tid == 0) {
        off[cs.nblk] = total;
        off[cs.nblk + 1] = total + nspill;
    }
    __syncthreads();
    if (tid == 0) st->spill = 0;  // ready for the next query
    const uint32_t npre = cs.pre ? st->pre_count : 0u;
    const uint32_t c = npre + off[cs.nblk + 1];
    const uint32_t bad = spilled > cs.spill_cap ? STATUS_OVERFLOW : (c < kreq ? STATUS_RETRY : 0u);
    if (bad) {
        if (tid == 0) {
            st->status |= bad;
            if (fo.status) *fo.status = bad;
        }
        return;
    }
    if (tid == 0) {
        st->cand_last = c;
        if (fo.cand) *fo.cand = c;
    }
    const SlotView v{cs.pre, npre, cs.keys, cs.slot, off, cs.nblk, cs.spill};
    select_and_emit<FINAL_BS, FINAL_KPT>(v, c, k, fo, top, sorted, ss);
}

// Fallback path for candidate sets too large for the regular slots: the
// per-block counts are scanned into global offsets, every block copies its
// slot region into one flat array, and the final select reads that array.
__global__ __launch_bounds__(FLAT_BS) void cand_offsets_kernel(FteState *__restrict__ st, CandSet cs,
                                                               uint32_t *__restrict__ offsets,
                                                               uint64_t *__restrict__ flat) {
    __shared__ uint32_t wsum[FLAT_BS / 64];
    __shared__ uint32_t total;
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (cs.nblk + FLAT_BS - 1) / FLAT_BS;
    const uint32_t b0 = tid * per < cs.nblk ? tid * per : cs.nblk;
    const uint32_t b1 = b0 + per < cs.nblk ? b0 + per : cs.nblk;
    uint32_t mine = 0;
    for (uint32_t b = b0; b < b1; b++) mine += cs.counts[b];
    const uint32_t npre = cs.pre ? st->pre_count : 0u;
    uint32_t run = npre + block_excl_scan<FLAT_BS>(mine, wsum, &total);
    for (uint32_t b = b0; b < b1; b++) {
        offsets[b] = run;
        run += cs.counts[b];
    }
    for (uint32_t i = tid; i < npre; i += FLAT_BS) flat[i] = cs.pre[i];
    if (tid == 0) {
        st->total = npre + total;
        st->spill = 0;
    }
}

__global__ __launch_bounds__(256) void cand_compact_kernel(CandSet cs, const uint32_t *__restrict__ offsets,
                                                           uint64_t *__restrict__ flat) {
    const uint32_t b = blockIdx.x;
    const uint32_t c = cs.counts[b];
    const uint64_t *src = cs.keys + (uint64_t)b * cs.slot;
    uint64_t *dst = flat + offsets[b];
    for (uint32_t i = threadIdx.x; i < c; i += 256) dst[i] = src[i];
}

__global__ __launch_bounds__(FLAT_BS) void fte_final_flat_kernel(FteState *__restrict__ st,
                                                                  const uint64_t *__restrict__ flat, uint32_t k,
                                                                  uint32_t kreq, FinalOut fo) {
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ uint64_t sorted[TOPK_MAX_K];
    __shared__ SelectScratch ss;
    const uint32_t c = st->total;
    if (c < kreq) {
        if (threadIdx.x == 0) {
            st->status |= STATUS_RETRY;
            if (fo.status) *fo.status = STATUS_RETRY;
        }
        return;
    }
    if (threadIdx.x == 0) st->cand_last = c;
    select_and_emit<FLAT_BS, 8>(FlatView{flat}, c, k, fo, top, sorted, ss);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static uint32_t exact_prefix_blocks(uint64_t n, uint32_t k) {
    const uint64_t need = (n + SEED_TILE - 1) / SEED_TILE;
    uint32_t g = 32768u / (k ? k : 1);
    if (g > 256) g = 256;
    if (g < 8) g = 8;
    return (uint32_t)(need < g ? need : g);
}

static uint32_t stream_grid(uint32_t grid_cap, uint64_t ntiles) {
    const uint64_t cap = grid_cap ? (uint64_t)grid_cap : 2048;
    const uint64_t g = ntiles < cap ? ntiles : cap;
    return (uint32_t)(g < 1 ? 1 : g);
}

// float4 vectors per lane per column of one stream tile: enough loads in
// flight per lane (about 8-16) without exhausting registers.
static int stream_vpt(uint32_t m) {
    if (m <= 1) return 8;
    if (m <= 4) return 4;
    if (m <= 8) return 2;
    return 1;
}

// Chooses the seed geometry for a query: sample mode when n is large enough
// that a spread sample leaves many candidates, exact prefix mode otherwise.
// Returns true for the sample geometry.
static bool seed_plan(uint64_t n, uint32_t k, bool exact, uint32_t &blocks, uint64_t &stride, uint64_t &n_eff,
                      uint32_t &ks, uint32_t &total_valid) {
    const uint64_t s = (uint64_t)SAMPLE_CHUNKS * SAMPLE_TILE;
    if (!exact && n > 4 * s) {
        blocks = SAMPLE_CHUNKS;
        stride = ((n / SAMPLE_CHUNKS) / SAMPLE_TILE) * SAMPLE_TILE;
        n_eff = n;
        // expected candidates ~ ks * n / s; aim for >= 8k
        const uint64_t want = (8ull * k * s + n - 1) / n;
        ks = (uint32_t)(want < 4 ? 4 : (want > k ? k : want));
        total_valid = blocks * ks;
        return true;
    } else {
        blocks = exact_prefix_blocks(n, k);
        stride = SEED_TILE;
        const uint64_t p = (uint64_t)blocks * SEED_TILE;
        n_eff = n < p ? n : p;
        ks = k;
        const uint64_t full = n_eff / SEED_TILE, rem = n_eff % SEED_TILE;
        total_valid = (uint32_t)(full * (k < SEED_TILE ? k : SEED_TILE) + (rem < k ? rem : k));
        return false;
    }
}

// Everything one run of the FTE pipeline needs besides the query: the store
// size, the stream, the scratch its kernels use and the launch shape.
struct FteRun {
    hipStream_t s;
    uint64_t n;
    FteState *st;
    uint64_t *blk;        // seed-pass block winners (SAMPLE_CHUNKS * TOPK_MAX_K keys)
    uint64_t *pre;        // exact-mode prefix keys (TOPK_MAX_K)
    uint32_t *counts;     // per-stream-block counts (MAX_STREAM_GRID)
    uint64_t *cand;       // candidate slots
    uint64_t cand_cap;
    bool big;             // cand holds every object of every block (exact re-run after an overflow)
    uint64_t *spill;
    uint32_t spill_cap;
    int vpt;              // 0: chosen from the query width
    uint32_t grid_cap;    // 0: default
    const uint32_t *gate; // non-null: every kernel exits at once unless *gate != 0
    topk_ctx *timed;      // non-null: bracket the stream pass for topk_timing_*
};

// Enqueues seed -> stream -> final select for one query on r.s.  Returns the
// candidate set the final select read (pta.hip re-uses the final stage).
static CandSet fte_launch(const FteRun &r, uint32_t k, const QCols &q, const FinalOut &fo, int mode) {
    const uint64_t n = r.n;
    const bool exact = mode != FTE_SAMPLE;
    hipStream_t s = r.s;

    uint32_t blocks, ks, total_valid;
    uint64_t stride, n_eff;
    const bool sampled = seed_plan(n, k, exact, blocks, stride, n_eff, ks, total_valid);
    uint64_t *pre = exact ? r.pre : nullptr;
    // sample mode with few winners: no merge in the seed pass, every stream
    // wave reduces the winners to tau itself
    const bool wave_tau = sampled && ks <= 4 && blocks * ks <= SEED_WAVE_MAX;
    // kernels unrolled for the query width, or the runtime-width ones (the
    // only ones that divide real-data scores)
    const uint32_t msw = q.div != 0.0f ? 0u : q.m;
#define TK_SEED(MM, OPT, MERGE)                                                                           \
    hipLaunchKernelGGL((fte_seed_kernel<MM, OPT, MERGE>), dim3(blocks), dim3(SEED_BS), 0, s, q, n_eff, stride, \
                       ks, total_valid, r.blk, pre, r.st, r.gate)
#define TK_SEED_SAMPLE(MM) TK_SEED(MM, SAMPLE_TILE / SEED_BS, true)
#define TK_SEED_EXACT(MM) TK_SEED(MM, SEED_TILE / SEED_BS, true)
#define TK_SAMPLE(MM) \
    hipLaunchKernelGGL(fte_sample_kernel<MM>, dim3(blocks), dim3(SEED_BS), 0, s, q, n_eff, stride, ks, r.blk)
    if (wave_tau) {
        TK_SWITCH_M(msw, TK_SAMPLE);
    } else if (sampled) {
        TK_SWITCH_M(msw, TK_SEED_SAMPLE);
    } else {
        TK_SWITCH_M(msw, TK_SEED_EXACT);
    }
#undef TK_SEED
#undef TK_SAMPLE
#undef TK_SEED_SAMPLE
#undef TK_SEED_EXACT
    const uint32_t nseed = wave_tau ? blocks * ks : 0u;

    // exact mode streams only the objects after the prefix
    const uint64_t first = exact ? n_eff : 0;
    const uint32_t kreq = (uint32_t)(n < k ? n : k);
    const int vpt = r.vpt ? r.vpt : stream_vpt(q.m);
    const uint64_t tile = (uint64_t)STREAM_BS * 4 * vpt;
    const uint64_t ntiles = n > first ? (n - first + tile - 1) / tile : 0;
    const uint32_t g = ntiles ? stream_grid(r.grid_cap, ntiles) : 0;
    CandSet cs;
    cs.keys = r.cand;
    cs.counts = r.counts;
    cs.nblk = g;
    cs.pre = pre;
    cs.spill = r.spill;
    cs.spill_cap = r.spill_cap;
    if (r.big) cs.slot = g ? (uint32_t)(((ntiles + g - 1) / g) * tile) : 0;  // a block's every object fits
    else cs.slot = g ? (uint32_t)(r.cand_cap / g) : 0;
    if (g) {
        hipEvent_t ev0 = nullptr;
        if (r.timed) timing_begin(r.timed, s, ev0);
#define TK_STREAM(MM, VPT)                                                                                   \
    hipLaunchKernelGGL((fte_stream_kernel<MM, VPT>), dim3(g), dim3(STREAM_BS), 0, s, q, n, first, ntiles, \
                       r.st, cs, r.blk, nseed, ks, r.gate)
#define TK_STREAM1(MM) TK_STREAM(MM, 1)
#define TK_STREAM2(MM) TK_STREAM(MM, 2)
#define TK_STREAM4(MM) TK_STREAM(MM, 4)
#define TK_STREAM8(MM) TK_STREAM(MM, 8)
        switch (vpt) {
            case 1: TK_SWITCH_M(msw, TK_STREAM1); break;
            case 2: TK_SWITCH_M(msw, TK_STREAM2); break;
            case 4: TK_SWITCH_M(msw, TK_STREAM4); break;
            default: TK_SWITCH_M(msw, TK_STREAM8); break;
        }
#undef TK_STREAM
#undef TK_STREAM1
#undef TK_STREAM2
#undef TK_STREAM4
#undef TK_STREAM8
        if (r.timed) timing_end(r.timed, s, "fte_stream", ev0);
    }
    if (!r.big) hipLaunchKernelGGL(fte_final_kernel, dim3(1), dim3(FINAL_BS), 0, s, r.st, cs, k, kreq, fo, r.gate);
    return cs;
}

int fte_enqueue(topk_ctx *ctx, Lane &L, uint32_t k, const QCols &q, uint64_t *d_ids, float *d_scores,
                float *d_thr, uint32_t *d_status, int mode, uint64_t n) {
    if (!n) n = ctx->n;
    if (n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;  // a run's ids are 32-bit: callers split larger stores
    FteRun r{};
    r.s = L.stream;
    r.n = n;
    r.st = L.st;
    r.blk = L.blk;
    r.pre = L.pre;
    r.counts = L.counts;
    r.big = mode == FTE_EXACT_BIG;
    r.cand = r.big ? ctx->cand_big : L.cand;
    r.cand_cap = ctx->cand_cap;
    r.spill = L.spill;
    r.spill_cap = ctx->spill_cap;
    r.vpt = ctx->tune.stream_vpt;
    r.grid_cap = ctx->tune.stream_grid;
    r.timed = ctx;
    const uint32_t kreq = (uint32_t)(n < k ? n : k);
    const FinalOut fo{d_ids, d_scores, d_thr, d_status};
    const CandSet cs = fte_launch(r, k, q, fo, mode);
    if (r.big) final_enqueue(ctx, L, cs, k, kreq, fo, true);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

// ---------------------------------------------------------------------------
// id shards: top k of the union of per-shard top-k lists
// ---------------------------------------------------------------------------
// Entry j of shard s's list (best first) lands at output position j plus,
// for every other shard, the number of its entries that rank above it -- a
// binary search, since each list is sorted.  Ranking is by score (monotone
// key) and then global id; shard t < s holds only lower ids, so on equal
// scores shard t's entries rank above shard s's.  Positions past k are
// dropped; the k-th entry's score is the threshold.
constexpr int MERGE_BS = 1024;
struct ShardCounts {
    uint32_t shards;
    uint32_t cnt[MAX_ID_SHARDS];  // entries of each shard's list: min(k, objects of the shard)
};

__global__ __launch_bounds__(MERGE_BS) void shard_merge_kernel(const uint64_t *__restrict__ sh_ids,
                                                               const float *__restrict__ sh_scores, ShardCounts c,
                                                               uint32_t k, uint32_t kreq, FinalOut fo) {
    const uint32_t S = c.shards;
    for (uint32_t e = threadIdx.x; e < S * k; e += MERGE_BS) {
        const uint32_t s = e / k, j = e % k;
        if (j >= c.cnt[s]) continue;
        const float sc = sh_scores[(uint64_t)s * TOPK_MAX_K + j];
        const uint32_t ok = okey(sc);
        uint32_t r = j;
        for (uint32_t t = 0; t < S && r < k; t++) {
            if (t == s) continue;
            const float *ts = sh_scores + (uint64_t)t * TOPK_MAX_K;
            uint32_t lo = 0, hi = c.cnt[t];  // entries [0, lo) rank above, [hi, cnt) below
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                const uint32_t om = okey(ts[mid]);
                if (om > ok || (om == ok && t < s)) lo = mid + 1;
                else hi = mid;
            }
            r += lo;
        }
        if (r < k) {  // positions are distinct: global ids are
            fo.ids[r] = (uint64_t)s * ID_SHARD + sh_ids[(uint64_t)s * TOPK_MAX_K + j];
            fo.scores[r] = sc;
            if (r == kreq - 1 && fo.thr) *fo.thr = sc;
        }
    }
    for (uint32_t i = kreq + threadIdx.x; i < k; i += MERGE_BS) {
        fo.ids[i] = ~0ull;
        fo.scores[i] = -INFINITY;
    }
    if (threadIdx.x == 0 && fo.status) *fo.status = 0;
}

void shard_merge_enqueue(hipStream_t s, const uint64_t *sh_ids, const float *sh_scores, uint64_t n, uint32_t shards,
                         uint32_t k, const FinalOut &fo) {
    ShardCounts c{};
    c.shards = shards;
    for (uint32_t i = 0; i < shards; i++) {
        const uint64_t ns = n - (uint64_t)i * ID_SHARD < ID_SHARD ? n - (uint64_t)i * ID_SHARD : ID_SHARD;
        c.cnt[i] = (uint32_t)(ns < k ? ns : k);
    }
    const uint32_t kreq = (uint32_t)(n < k ? n : k);
    hipLaunchKernelGGL(shard_merge_kernel, dim3(1), dim3(MERGE_BS), 0, s, sh_ids, sh_scores, c, k, kreq, fo);
}

void final_enqueue(topk_ctx *ctx, Lane &L, const CandSet &cs, uint32_t k, uint32_t kreq, const FinalOut &fo,
                   bool flat_fallback) {
    hipStream_t s = L.stream;
    if (flat_fallback) {
        hipLaunchKernelGGL(cand_offsets_kernel, dim3(1), dim3(FLAT_BS), 0, s, L.st, cs, ctx->offsets, ctx->flat);
        if (cs.nblk)
            hipLaunchKernelGGL(cand_compact_kernel, dim3(cs.nblk), dim3(256), 0, s, cs, ctx->offsets, ctx->flat);
        hipLaunchKernelGGL(fte_final_flat_kernel, dim3(1), dim3(FLAT_BS), 0, s, L.st, ctx->flat, k, kreq, fo);
    } else {
        hipLaunchKernelGGL(fte_final_kernel, dim3(1), dim3(FINAL_BS), 0, s, L.st, cs, k, kreq, fo, nullptr);
    }
}

// ---------------------------------------------------------------------------
// allocation-free FTE over a caller's device store (topk_fte_noalloc)
// ---------------------------------------------------------------------------
// The per-query scratch lives in one file-scope device array per GPU, held
// until the process ends: FTE state, the two passes' status words, seed
// winners, prefix keys, per-block counts, candidate slots and spill region.
// A query is at most two pipeline runs, both enqueued up front: the first
// (sample-seeded above 2^19 objects, exact-prefix-seeded below) and an
// exact-prefix run whose kernels exit at once unless the first reported a
// failure.  Nothing waits on the host, so the call can be captured in a graph.
constexpr uint64_t NA_CAND = 1ull << 20;   // candidate slots (8 MiB)
constexpr uint32_t NA_SPILL = 1u << 18;    // spill region (2 MiB)
constexpr uint64_t NA_BLK = (uint64_t)SAMPLE_CHUNKS * TOPK_MAX_K;  // seed winners, either mode (1 MiB)
struct NoAllocScratch {
    FteState st;
    uint32_t pass[2];  // final-select status of the first and the second run (0: settled)
    uint32_t counts[MAX_STREAM_GRID];
    uint64_t pre[TOPK_MAX_K];
    uint64_t blk[NA_BLK];
    uint64_t cand[NA_CAND];
    uint64_t spill[NA_SPILL];
};
static_assert(offsetof(NoAllocScratch, pre) % 8 == 0 && offsetof(NoAllocScratch, blk) % 8 == 0 &&
                  offsetof(NoAllocScratch, cand) % 8 == 0 && offsetof(NoAllocScratch, spill) % 8 == 0,
              "key arrays stay 8-byte aligned");
constexpr size_t NA_SCRATCH_BYTES = sizeof(NoAllocScratch);

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

// Device address of `scratch` per GPU (the array's name on the host is only
// a stand-in for the symbol), looked up once per device.
constexpr int NA_MAX_DEVICES = 64;
static std::mutex na_mu;
static NoAllocScratch *na_ptr[NA_MAX_DEVICES];

static int noalloc_scratch(NoAllocScratch **out) {
    int dev = 0;
    TK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return TOPK_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        TK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return TOPK_ERR_HIP;
        na_ptr[dev] = static_cast<NoAllocScratch *>(p);
    }
    *out = na_ptr[dev];
    return TOPK_OK;
}

// Clears what the pipeline expects at zero (a kernel, not a memset node).
__global__ void noalloc_clear_kernel(NoAllocScratch *__restrict__ sc) {
    if (threadIdx.x == 0) {
        sc->st = FteState{};
        sc->pass[0] = 0;
        sc->pass[1] = 0;
    }
}

// Caller-visible status: TOPK_OK when either run settled the query.
__global__ void noalloc_status_kernel(const NoAllocScratch *__restrict__ sc, bool second,
                                      uint32_t *__restrict__ status) {
    if (threadIdx.x != 0) return;
    const uint32_t last = second && sc->pass[0] ? sc->pass[1] : sc->pass[0];
    *status = last == 0 ? (uint32_t)TOPK_OK : (uint32_t)TOPK_ERR_OVERFLOW;
}

}  // namespace topk

using namespace topk;

extern "C" int topk_fte_noalloc(const float *d_cols, uint64_t n, uint32_t d, uint64_t ld, uint32_t k, uint8_t qq,
                                const float *weights, const uint8_t *attr, uint64_t *d_ids, float *d_scores,
                                float *d_threshold, uint32_t *d_status, void *stream) {
    if (!d_cols || !weights || !attr || !d_ids || !d_scores || !d_status) return TOPK_ERR_INVALID;
    if (n < 1 || d < 1 || d > TOPK_MAX_DIMS || ld < n || ld % 4 || k < 1) return TOPK_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(d_cols) % 16) return TOPK_ERR_INVALID;
    if (n >= 0xFFFFFFFFull || k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    if (qq < 1 || qq > d) return TOPK_ERR_INVALID;
    QCols q;
    memset(&q, 0, sizeof(q));
    for (uint32_t j = 0; j < qq; j++) {
        if (attr[j] >= d) return TOPK_ERR_INVALID;
        const float wj = weights[attr[j]];
        if (!std::isfinite(wj)) return TOPK_ERR_INVALID;
        q.col[j] = d_cols + (uint64_t)attr[j] * ld;
        q.w[j] = wj;
    }
    q.m = qq;
    NoAllocScratch *sc = nullptr;
    int rc = noalloc_scratch(&sc);
    if (rc) return rc;

    FteRun r{};
    r.s = static_cast<hipStream_t>(stream);
    r.n = n;
    r.st = &sc->st;
    r.blk = sc->blk;
    r.pre = sc->pre;
    r.counts = sc->counts;
    r.cand = sc->cand;
    r.cand_cap = NA_CAND;
    r.spill = sc->spill;
    r.spill_cap = NA_SPILL;
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(1), dim3(64), 0, r.s, sc);
    uint32_t blocks, ks, total_valid;
    uint64_t stride, n_eff;
    const bool sampled = seed_plan(n, k, false, blocks, stride, n_eff, ks, total_valid);
    fte_launch(r, k, q, FinalOut{d_ids, d_scores, d_threshold, &sc->pass[0]}, FTE_SAMPLE);
    if (sampled) {
        // too few candidates (the sample guessed high) or a full buffer: the
        // exact-prefix run, skipped on the device when the first one settled
        r.gate = &sc->pass[0];
        fte_launch(r, k, q, FinalOut{d_ids, d_scores, d_threshold, &sc->pass[1]}, FTE_EXACT);
    }
    hipLaunchKernelGGL(noalloc_status_kernel, dim3(1), dim3(64), 0, r.s, sc, sampled, d_status);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}
