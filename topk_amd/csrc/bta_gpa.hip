// Device-wide radix select, bitonic top-k (BTA) and iterative prune+compact
// (GPA) on gfx950.  All three return what topk_fte returns for the same
// query (radix select: the element of a given rank).
//
//  * radix select (gpu/radix_select.h:73-120): the reference runs eight
//    4-bit passes, each followed by a host sync.  Here: three passes of
//    11 / 11 / 10-bit digits; each pass builds its digit histogram in LDS
//    (one global atomic per non-empty bin per workgroup) and a one-wave
//    kernel picks the digit on the device, so a select is six launches and
//    no host round trip.
//  * BTA (gpu/BTA.h:57-73 local_sort2, :124-167 local_sort3, :207-255): a
//    workgroup scores BTA_OBJS (2048) objects into LDS keys (the reference
//    blocks 4096), bitonic-sorts them into
//    runs of K2 = next_pow2(k) with alternating direction, then halves the
//    run count with the max-of-pairs step (a descending run against an
//    ascending one keeps the K2 largest as a bitonic sequence) and a bitonic
//    clean, until one run remains.  Block runs are merged the same way by
//    tree launches of BTA_OBJS / K2 runs per workgroup.  Keys carry ids (the
//    reference keeps scores only, gpu/BTA.h:17).
//  * GPA (gpu/GPA.h:161-227, gpu/prune_tuples.h:10-118): tuples (id,
//    partial score) over the query's terms in order.  After term j the k-th
//    largest partial score s_k is found by radix select; every final score
//    lies within [s + lo_j - e(s), s + hi_j + e(s)], where lo_j / hi_j sum
//    the remaining terms' extreme contributions from per-column min / max
//    and e covers binary32 rounding of the remaining steps, so
//    thr = s_k + lo_j - e(s_k) is a lower bound of the final k-th score and a
//    tuple with s + hi_j + e(s) < thr cannot be in the top k.  Survivors are
//    compacted with the next term added.  The reference's bound
//    (score + x_C * (d-1-it), gpu/prune_tuples.h:14) assumes the reordered
//    store, all attributes and unit weights; this one holds for any weights
//    and attribute order.
#include <math.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "device.hpp"
#include "emit.hpp"
#include "internal.hpp"

namespace topk {

// ---------------------------------------------------------------------------
// device-wide radix select
// ---------------------------------------------------------------------------
constexpr int DSEL_BS = 256;
constexpr uint32_t DSEL_MAX_GRID = 2048;
constexpr int DSEL_PASSES = 3;
constexpr int DSEL_SHIFT[DSEL_PASSES] = {21, 10, 0};
constexpr int DSEL_WIDTH[DSEL_PASSES] = {11, 11, 10};

struct DselState {
    uint32_t prefix, mask, rank, pad;
    uint32_t hist[2048];
};

// Key sources: one key, or keys 4q..4q+3 with one 16-byte load (the base
// pointer is 16-byte aligned when `quads` is used).
struct U32Src {
    const uint32_t *k;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return k[i]; }
    __device__ __forceinline__ uint4 quad(uint64_t q) const { return reinterpret_cast<const uint4 *>(k)[q]; }
};
struct ScoreSrc {
    const float *s;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return okey(s[i]); }
    __device__ __forceinline__ uint4 quad(uint64_t q) const {
        const float4 v = reinterpret_cast<const float4 *>(s)[q];
        return make_uint4(okey(v.x), okey(v.y), okey(v.z), okey(v.w));
    }
};

// rank = (cnt ? *cnt : base) - sub
__global__ void dsel_init_kernel(DselState *st, const uint32_t *cnt, uint64_t base, uint64_t sub) {
    for (uint32_t i = threadIdx.x; i < 2048; i += blockDim.x) st->hist[i] = 0;
    if (threadIdx.x == 0) {
        const uint64_t n = cnt ? *cnt : base;
        st->prefix = 0;
        st->mask = 0;
        st->rank = (uint32_t)(n - sub);
    }
}

// Digit histogram of one pass in LDS.  The keys of a wave are folded with
// wave primitives before they reach LDS: the lanes sharing the first
// pending lane's digit (a ballot of digit equality) add their count with
// one atomic from that lane, twice per key slot, and the rest add one by
// one.  Keys with concentrated digits -- score keys share their exponent
// bits, and a late pass sees one bucket's keys -- then cost one or two LDS
// atomics per wave instead of up to 64 on one address; spread digits cost
// two extra ballots.  Every lane runs the same trip count (ballots inside).
constexpr int DSEL_FOLD = 2;  // leader-digit folds per key slot
template <class Src>
__global__ __launch_bounds__(DSEL_BS) void dsel_hist_kernel(Src src, const uint32_t *cnt, uint64_t n_host,
                                                            DselState *st, int shift, int width, int quads) {
    __shared__ uint32_t h[2048];
    const uint32_t nb = 1u << width, lane = lane_id();
    for (uint32_t i = threadIdx.x; i < nb; i += DSEL_BS) h[i] = 0;
    __syncthreads();
    const uint64_t n = cnt ? *cnt : n_host;
    const uint32_t prefix = st->prefix, mask = st->mask;
    auto take = [&](uint32_t k, bool valid) {
        const uint32_t d = (k >> shift) & (nb - 1);
        bool pend = valid && (k & mask) == prefix;
#pragma unroll
        for (int it = 0; it < DSEL_FOLD; it++) {
            const uint64_t act = __ballot(pend);
            if (!act) break;
            const uint32_t lead = (uint32_t)__builtin_ctzll(act);
            const uint32_t ld = (uint32_t)__shfl((int)d, (int)lead, 64);
            const bool same = pend && d == ld;
            const uint64_t sm = __ballot(same);
            if (lane == lead) atomicAdd(&h[ld], (uint32_t)__popcll(sm));
            pend = pend && !same;
        }
        if (pend) atomicAdd(&h[d], 1u);
    };
    const uint64_t gid = (uint64_t)blockIdx.x * DSEL_BS + threadIdx.x, stride = (uint64_t)gridDim.x * DSEL_BS;
    const uint64_t n4 = quads ? n / 4 : 0;
    for (uint64_t qb = gid - lane; qb < n4; qb += stride) {  // wave-uniform trip count
        const uint64_t q = qb + lane;
        const bool valid = q < n4;
        const uint4 k = valid ? src.quad(q) : make_uint4(0u, 0u, 0u, 0u);
        take(k.x, valid);
        take(k.y, valid);
        take(k.z, valid);
        take(k.w, valid);
    }
    for (uint64_t ib = n4 * 4 + gid - lane; ib < n; ib += stride) {
        const uint64_t i = ib + lane;
        const bool valid = i < n;
        take(valid ? src(i) : 0u, valid);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += DSEL_BS)
        if (h[i]) atomicAdd(&st->hist[i], h[i]);
}

// One wave: the smallest digit whose ascending cumulative count exceeds the
// remaining rank; lane l owns bins [l*per, (l+1)*per).
__global__ __launch_bounds__(64) void dsel_pick_kernel(DselState *st, int shift, int width) {
    const uint32_t nb = 1u << width, per = nb / 64, lane = threadIdx.x;
    const uint32_t rank = st->rank;
    uint32_t c[32];
    uint32_t tot = 0;
    for (uint32_t j = 0; j < per; j++) {
        c[j] = st->hist[lane * per + j];
        tot += c[j];
    }
    const uint32_t incl = wave_incl_scan(tot), excl = incl - tot;
    if (excl <= rank && rank < incl) {
        uint32_t acc = excl;
        for (uint32_t j = 0; j < per; j++) {
            if (rank < acc + c[j]) {
                st->prefix |= (lane * per + j) << shift;
                st->mask |= (nb - 1) << shift;
                st->rank = rank - acc;
                break;
            }
            acc += c[j];
        }
    }
    for (uint32_t j = 0; j < per; j++) st->hist[lane * per + j] = 0;
}

template <class Src>
static void dsel_enqueue(hipStream_t s, Src src, const uint32_t *cnt, uint64_t n_bound, uint64_t base, uint64_t sub,
                         DselState *st, bool aligned) {
    uint32_t g = (uint32_t)((n_bound / (aligned ? 4 : 1) + DSEL_BS - 1) / DSEL_BS);
    if (g > DSEL_MAX_GRID) g = DSEL_MAX_GRID;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(dsel_init_kernel, dim3(1), dim3(256), 0, s, st, cnt, base, sub);
    for (int p = 0; p < DSEL_PASSES; p++) {
        hipLaunchKernelGGL((dsel_hist_kernel<Src>), dim3(g), dim3(DSEL_BS), 0, s, src, cnt, n_bound, st,
                           DSEL_SHIFT[p], DSEL_WIDTH[p], aligned ? 1 : 0);
        hipLaunchKernelGGL(dsel_pick_kernel, dim3(1), dim3(64), 0, s, st, DSEL_SHIFT[p], DSEL_WIDTH[p]);
    }
}

// ---------------------------------------------------------------------------
// BTA
// ---------------------------------------------------------------------------
constexpr int BTA_BS = 256;
constexpr uint32_t BTA_OBJS = 2048;  // objects per block (32 KiB of LDS per workgroup: 4 workgroups per CU)
static_assert(BTA_OBJS >= 2 * TOPK_MAX_K, "a merge workgroup combines at least two runs of up to TOPK_MAX_K keys");

__device__ __forceinline__ void cmpx(uint64_t *a, uint32_t lo, uint32_t hi, bool desc) {
    const uint64_t x = a[lo], y = a[hi];
    if ((x < y) == desc) {
        a[lo] = y;
        a[hi] = x;
    }
}

// Bitonic sort of every run of K2 keys of a[0..BTA_OBJS); run r ends descending
// when r is even and ascending when r is odd.
__device__ void bta_sort_runs(uint64_t *a, uint32_t K2) {
    for (uint32_t size = 2; size <= K2; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < BTA_OBJS / 2; t += BTA_BS) {
                const uint32_t lo = 2 * stride * (t / stride) + (t & (stride - 1));
                cmpx(a, lo, lo + stride, (lo & size) == 0);
            }
            __syncthreads();
        }
}

// Halves nruns alternating runs of K2 (even descending, odd ascending) until
// one descending run holds the K2 largest keys; returns the buffer holding it.
__device__ uint64_t *bta_merge_runs(uint64_t *src, uint64_t *dst, uint32_t nruns, uint32_t K2) {
    while (nruns > 1) {
        const uint32_t half = nruns / 2;
        for (uint32_t t = threadIdx.x; t < half * K2; t += BTA_BS) {
            const uint32_t r = t / K2, i = t - r * K2;
            const uint64_t x = src[2 * r * K2 + i], y = src[(2 * r + 1) * K2 + i];
            dst[t] = x > y ? x : y;
        }
        __syncthreads();
        const uint32_t hk = K2 / 2;
        for (uint32_t stride = hk; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < half * hk; t += BTA_BS) {
                const uint32_t r = t / hk, u = t - r * hk;
                const uint32_t lo = r * K2 + 2 * stride * (u / stride) + (u & (stride - 1));
                cmpx(dst, lo, lo + stride, (r & 1) == 0);
            }
            __syncthreads();
        }
        uint64_t *tmp = src;
        src = dst;
        dst = tmp;
        nruns = half;
    }
    return src;
}

// One workgroup per BTA_OBJS objects: scores -> keys -> the block's K2 best.
__global__ __launch_bounds__(BTA_BS) void bta_block_kernel(QCols q, uint64_t n, uint32_t K2,
                                                           uint64_t *__restrict__ runs) {
    __shared__ uint64_t a[BTA_OBJS], b[BTA_OBJS];
    const uint64_t base = (uint64_t)blockIdx.x * BTA_OBJS;
#pragma unroll
    for (int j = 0; j < (int)(BTA_OBJS / (4 * BTA_BS)); j++) {
        const uint32_t o = j * 4 * BTA_BS + threadIdx.x * 4;
        const uint64_t v = (base + o) / 4;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t t = 0; t < q.m; t++) {
            const f4 x = *reinterpret_cast<const f4 *>(q.col[t] + v * 4);
            const f4 p = x * q.w[t];
            acc = acc + p;
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const uint64_t id = base + o + e;
            a[o + e] = id < n ? rkey(acc[e], (uint32_t)id) : 0ull;
        }
    }
    __syncthreads();
    bta_sort_runs(a, K2);
    const uint64_t *r = bta_merge_runs(a, b, BTA_OBJS / K2, K2);
    for (uint32_t i = threadIdx.x; i < K2; i += BTA_BS) runs[(uint64_t)blockIdx.x * K2 + i] = r[i];
}

// ---- register-resident block kernel for runs of 64 keys (k <= 64) ----
// The key of lane l's partner lane l ^ S, each 32-bit half moved without
// LDS where a lane pattern allows: quad permutes (S = 1, 2), a row rotation
// by 8 (S = 8: lane l ^ 8 either way round), swizzles inside 32-lane halves
// (S = 4, 16) and a permute for S = 32.
template <int S>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
    if constexpr (S == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (S == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (S == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
    else if constexpr (S == 4 || S == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (S << 10));
    else return (uint32_t)__shfl_xor((int)v, S, 64);
}

template <int S>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v) {
    const uint32_t lo = xor_lane<S>((uint32_t)v), hi = xor_lane<S>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Compare-exchange with lane l ^ S: the lane keeps the smaller key when
// keep_min, else the larger.
template <int S>
__device__ __forceinline__ void cx_lane(uint64_t &v, bool keep_min) {
    const uint64_t p = xor_lane64<S>(v);
    v = (p < v) == keep_min ? p : v;
}

// Bitonic merge of a 64-lane bitonic sequence into descending (desc) or
// ascending lane order.
__device__ __forceinline__ void clean64(uint64_t &v, uint32_t l, bool desc) {
    cx_lane<32>(v, ((l & 32) == 0) != desc);
    cx_lane<16>(v, ((l & 16) == 0) != desc);
    cx_lane<8>(v, ((l & 8) == 0) != desc);
    cx_lane<4>(v, ((l & 4) == 0) != desc);
    cx_lane<2>(v, ((l & 2) == 0) != desc);
    cx_lane<1>(v, ((l & 1) == 0) != desc);
}

// Bitonic sort of one key per lane into descending or ascending lane order.
__device__ __forceinline__ void sort64(uint64_t &v, uint32_t l, bool desc) {
    // stage (size, stride): the lane keeps the minimum iff its block of
    // `size` ascends and it is the lower lane of its pair (or neither)
    cx_lane<1>(v, (((l & 2) == 0) == ((l & 1) == 0)) != desc);
    cx_lane<2>(v, (((l & 4) == 0) == ((l & 2) == 0)) != desc);
    cx_lane<1>(v, (((l & 4) == 0) == ((l & 1) == 0)) != desc);
    cx_lane<4>(v, (((l & 8) == 0) == ((l & 4) == 0)) != desc);
    cx_lane<2>(v, (((l & 8) == 0) == ((l & 2) == 0)) != desc);
    cx_lane<1>(v, (((l & 8) == 0) == ((l & 1) == 0)) != desc);
    cx_lane<8>(v, (((l & 16) == 0) == ((l & 8) == 0)) != desc);
    cx_lane<4>(v, (((l & 16) == 0) == ((l & 4) == 0)) != desc);
    cx_lane<2>(v, (((l & 16) == 0) == ((l & 2) == 0)) != desc);
    cx_lane<1>(v, (((l & 16) == 0) == ((l & 1) == 0)) != desc);
    cx_lane<16>(v, (((l & 32) == 0) == ((l & 16) == 0)) != desc);
    cx_lane<8>(v, (((l & 32) == 0) == ((l & 8) == 0)) != desc);
    cx_lane<4>(v, (((l & 32) == 0) == ((l & 4) == 0)) != desc);
    cx_lane<2>(v, (((l & 32) == 0) == ((l & 2) == 0)) != desc);
    cx_lane<1>(v, (((l & 32) == 0) == ((l & 1) == 0)) != desc);
    clean64(v, l, desc);  // size 64: one block, ascending unless desc
}

// One workgroup per BTA_OBJS objects, K2 = 64: wave w scores its 512
// objects into 8 keys per lane, sorts each register across the 64 lanes
// into a run (even registers descending, odd ascending), halves the runs in
// registers (lane-wise max of a descending and an ascending run is a
// bitonic sequence holding their 64 largest keys, then a clean) down to
// one run per wave, and the four waves' runs are halved twice through LDS.
__global__ __launch_bounds__(BTA_BS) void bta_block64_kernel(QCols q, uint64_t n, uint64_t *__restrict__ runs) {
    __shared__ uint64_t x[BTA_BS / 64][64];
    const uint32_t l = lane_id(), w = threadIdx.x / 64;
    const uint64_t wbase = (uint64_t)blockIdx.x * BTA_OBJS + (uint64_t)w * 512;
    uint64_t key[8];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint64_t o = wbase + (uint64_t)h * 256 + (uint64_t)l * 4;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t t = 0; t < q.m; t++) {
            const f4 xv = *reinterpret_cast<const f4 *>(q.col[t] + o);
            const f4 p = xv * q.w[t];
            acc = acc + p;
        }
#pragma unroll
        for (int e = 0; e < 4; e++) key[4 * h + e] = o + e < n ? rkey(acc[e], (uint32_t)(o + e)) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) sort64(key[j], l, (j & 1) == 0);
#pragma unroll
    for (int p = 0; p < 4; p++) {
        key[p] = key[2 * p] > key[2 * p + 1] ? key[2 * p] : key[2 * p + 1];
        clean64(key[p], l, (p & 1) == 0);
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
        key[p] = key[2 * p] > key[2 * p + 1] ? key[2 * p] : key[2 * p + 1];
        clean64(key[p], l, (p & 1) == 0);
    }
    key[0] = key[0] > key[1] ? key[0] : key[1];
    clean64(key[0], l, (w & 1) == 0);  // odd waves ascend for the next halving
    x[w][l] = key[0];
    __syncthreads();
    if ((w & 1) == 0) {
        uint64_t v = x[w][l] > x[w + 1][l] ? x[w][l] : x[w + 1][l];
        clean64(v, l, w == 0);  // wave 0's run descends, wave 2's ascends
        key[0] = v;
    }
    __syncthreads();
    if ((w & 1) == 0) x[w][l] = key[0];
    __syncthreads();
    if (w == 0) {
        uint64_t v = x[0][l] > x[2][l] ? x[0][l] : x[2][l];
        clean64(v, l, true);
        runs[(uint64_t)blockIdx.x * 64 + l] = v;
    }
}

// Workgroup g merges runs [g*G, g*G + G) of `in` (descending, K2 keys each;
// missing runs are empty) into one descending run of `out`.
__global__ __launch_bounds__(BTA_BS) void bta_merge_kernel(const uint64_t *__restrict__ in, uint32_t nruns,
                                                           uint32_t K2, uint64_t *__restrict__ out) {
    __shared__ uint64_t a[BTA_OBJS], b[BTA_OBJS];
    const uint32_t G = BTA_OBJS / K2;
    const uint32_t r0 = blockIdx.x * G;
    for (uint32_t t = threadIdx.x; t < BTA_OBJS; t += BTA_BS) {
        const uint32_t lr = t / K2, i = t - lr * K2;
        const uint32_t r = r0 + lr;
        const uint32_t src_i = (lr & 1) ? K2 - 1 - i : i;  // odd runs ascending
        a[t] = r < nruns ? in[(uint64_t)r * K2 + src_i] : 0ull;
    }
    __syncthreads();
    const uint64_t *res = bta_merge_runs(a, b, G, K2);
    for (uint32_t i = threadIdx.x; i < K2; i += BTA_BS) out[(uint64_t)blockIdx.x * K2 + i] = res[i];
}

__global__ void emit_run_kernel(const uint64_t *__restrict__ run, uint32_t c, uint32_t k, uint64_t *__restrict__ ids,
                                float *__restrict__ scores, float *__restrict__ thr, uint32_t *__restrict__ status) {
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
        if (i < c) {
            ids[i] = rkey_id(run[i]);
            scores[i] = rkey_score(run[i]);
        } else {
            ids[i] = ~0ull;
            scores[i] = -INFINITY;
        }
    }
    if (threadIdx.x == 0) {
        if (thr) *thr = c ? rkey_score(run[c - 1]) : -INFINITY;
        if (status) *status = 0;
    }
}

// ---------------------------------------------------------------------------
// GPA
// ---------------------------------------------------------------------------
constexpr int GPA_BS = 256;
constexpr uint32_t GPA_MAX_GRID = 2048;

__global__ __launch_bounds__(GPA_BS) void gpa_first_kernel(const float *__restrict__ col, float w, uint64_t n,
                                                           uint32_t *__restrict__ ids, float *__restrict__ s,
                                                           uint32_t *__restrict__ cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * GPA_BS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GPA_BS) {
        const float p = col[i] * w;
        ids[i] = (uint32_t)i;
        s[i] = 0.0f + p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *cnt = (uint32_t)n;
}

// Remaining-term constants after the current term (FP64).
struct GpaStep {
    double hi, lo, r;   // sum of max / min contributions, sum of |w| max|x|
    int last;           // no terms remain: bounds are exact
};

__device__ __forceinline__ double gpa_err(double s, double r) { return 1.52587890625e-05 * (fabs(s) + r) + 1e-30; }

// thr = s_k + lo - e(s_k), rounded down (s_k = selected partial score).
__global__ void gpa_thr_kernel(const DselState *st, GpaStep g, float *thr) {
    const float sk = okey_inv(st->prefix);
    if (g.last) {
        *thr = sk;
        return;
    }
    const double t = (double)sk + g.lo - gpa_err(sk, g.r);
    float f = (float)t;
    if ((double)f > t) f = nextafterf(f, -INFINITY);
    *thr = f;
}

// Keeps tuples whose upper bound reaches thr, adding the next term (col, w)
// to the survivors' partial scores (col == nullptr: keep the score).
// Stable compaction in two kernels over the same grid: workgroup b owns the
// contiguous range [b*per, (b+1)*per) of the current survivors; the first
// kernel counts what it keeps, the second writes its survivors at the sum of
// the earlier workgroups' counts, in order (no global atomics).
__device__ __forceinline__ bool gpa_keep(float s, const GpaStep &g, float thr) {
    return g.last ? s >= thr : (double)s + g.hi + gpa_err(s, g.r) >= (double)thr;
}

__device__ __forceinline__ void gpa_range(uint32_t n, uint64_t &b, uint64_t &e) {
    const uint64_t per = ((uint64_t)n + gridDim.x - 1) / gridDim.x;
    b = (uint64_t)blockIdx.x * per;
    e = b + per < n ? b + per : n;
}

__global__ __launch_bounds__(GPA_BS) void gpa_count_kernel(const float *__restrict__ s_in,
                                                           const uint32_t *__restrict__ cnt_in, GpaStep g,
                                                           const float *__restrict__ thr_p,
                                                           uint32_t *__restrict__ bcount) {
    const uint32_t n = *cnt_in;
    const float thr = *thr_p;
    uint64_t b, e;
    gpa_range(n, b, e);
    uint32_t c = 0;
    for (uint64_t i = b + threadIdx.x; i < e; i += GPA_BS) c += gpa_keep(s_in[i], g, thr) ? 1u : 0u;
    __shared__ uint32_t ws[GPA_BS / 64];
    c = wave_sum_u32(c);
    if (lane_id() == 0) ws[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < GPA_BS / 64; w++) t += ws[w];
        bcount[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(GPA_BS) void gpa_prune_kernel(const uint32_t *__restrict__ ids_in,
                                                           const float *__restrict__ s_in,
                                                           const uint32_t *__restrict__ cnt_in, GpaStep g,
                                                           const float *__restrict__ thr_p,
                                                           const uint32_t *__restrict__ bcount,
                                                           const float *__restrict__ col, float w,
                                                           uint32_t *__restrict__ ids_out, float *__restrict__ s_out,
                                                           uint32_t *__restrict__ cnt_out) {
    const uint32_t n = *cnt_in;
    const float thr = *thr_p;
    uint64_t b, e;
    gpa_range(n, b, e);
    __shared__ uint32_t wsum[GPA_BS / 64];
    __shared__ uint32_t total;
    // this workgroup's output offset: the earlier workgroups' counts
    uint32_t part = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += GPA_BS) part += bcount[j];
    block_excl_scan<GPA_BS>(part, wsum, &total);
    uint32_t out = total;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *cnt_out = out + bcount[blockIdx.x];
    for (uint64_t i0 = b; i0 < e; i0 += GPA_BS) {
        const uint64_t i = i0 + threadIdx.x;
        bool keep = false;
        float sc = 0.0f;
        uint32_t id = 0;
        if (i < e) {
            sc = s_in[i];
            id = ids_in[i];
            keep = gpa_keep(sc, g, thr);
        }
        const uint32_t ex = block_excl_scan<GPA_BS>(keep ? 1u : 0u, wsum, &total);
        if (keep) {
            float ns = sc;
            if (col) {
                const float p = col[id] * w;
                ns = sc + p;
            }
            ids_out[out + ex] = id;
            s_out[out + ex] = ns;
        }
        out += total;
    }
}

struct GpaView {
    const float *s;
    const uint32_t *ids;
    __device__ __forceinline__ uint64_t get(uint32_t i) const { return rkey(s[i], ids[i]); }
};

__global__ __launch_bounds__(512) void gpa_final_kernel(const float *__restrict__ s, const uint32_t *__restrict__ ids,
                                                        const uint32_t *__restrict__ cnt, uint32_t k, FinalOut fo) {
    __shared__ uint64_t top[TOPK_MAX_K];
    __shared__ uint64_t sorted[TOPK_MAX_K];
    __shared__ SelectScratch ss;
    select_and_emit<512, 16>(GpaView{s, ids}, *cnt, k, fo, top, sorted, ss);
}

__global__ __launch_bounds__(256) void colstats_kernel(const float *__restrict__ cols, uint64_t n, uint64_t ld,
                                                       uint32_t *__restrict__ kmax, uint32_t *__restrict__ kmin) {
    const uint32_t m = blockIdx.y;
    const float *c = cols + (uint64_t)m * ld;
    float mx = -INFINITY, mn = INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const float v = c[i];
        mx = v > mx ? v : mx;
        mn = v < mn ? v : mn;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float a = __shfl_xor(mx, o, 64), b = __shfl_xor(mn, o, 64);
        mx = a > mx ? a : mx;
        mn = b < mn ? b : mn;
    }
    if (lane_id() == 0) {
        atomicMax(&kmax[m], okey(mx));
        atomicMin(&kmin[m], okey(mn));
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
struct AuxState {
    DselState *dsel = nullptr;
    uint64_t *bta[2] = {nullptr, nullptr};
    uint64_t bta_cap = 0;
    uint32_t *gids[2] = {nullptr, nullptr};
    float *gs[2] = {nullptr, nullptr};
    uint32_t *gcnt = nullptr;   // [2] survivor counts
    float *gthr = nullptr;
    uint32_t *gbcount = nullptr;  // survivors kept per compaction workgroup
    uint64_t gcap = 0;
    uint32_t *cstat = nullptr;  // okeys: max[d] | min[d]
    std::vector<float> cmax, cmin;
    const float *cstat_cols = nullptr;  // store the column statistics were taken from
    uint64_t cstat_gen = ~0ull;
};

static AuxState &aux(topk_ctx *ctx) {
    if (!ctx->aux) ctx->aux = new AuxState();
    return *static_cast<AuxState *>(ctx->aux);
}

void aux_free(topk_ctx *ctx) {
    AuxState *A = static_cast<AuxState *>(ctx->aux);
    if (!A) return;
    hipFree(A->dsel);
    hipFree(A->bta[0]);
    hipFree(A->bta[1]);
    hipFree(A->gids[0]);
    hipFree(A->gids[1]);
    hipFree(A->gs[0]);
    hipFree(A->gs[1]);
    hipFree(A->gcnt);
    hipFree(A->gthr);
    hipFree(A->gbcount);
    hipFree(A->cstat);
    delete A;
    ctx->aux = nullptr;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int ensure_dsel(AuxState &A) {
    if (!A.dsel) TK_HIP(hipMalloc(&A.dsel, sizeof(DselState)));
    return TOPK_OK;
}

// Copies a final result block (ctx->d_res) to the caller after sync.
static int finish_single(topk_ctx *ctx, hipStream_t s, uint32_t k, uint64_t *out_ids, float *out_scores,
                         float *out_thr) {
    TK_HIP(hipMemcpyAsync(ctx->h_res, ctx->d_res, ctx->res.bytes, hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    const uint64_t *ids = reinterpret_cast<const uint64_t *>(ctx->h_res + ctx->res.ids_off);
    const float *sc = reinterpret_cast<const float *>(ctx->h_res + ctx->res.scores_off);
    if (out_ids) memcpy(out_ids, ids, sizeof(uint64_t) * k);
    if (out_scores) memcpy(out_scores, sc, sizeof(float) * k);
    if (out_thr) *out_thr = *reinterpret_cast<const float *>(ctx->h_res + ctx->res.thr_off);
    return TOPK_OK;
}

// Per-column max / min of the store, cached until the store changes.
static int column_stats(topk_ctx *ctx, AuxState &A) {
    if (A.cstat_cols == ctx->cols && A.cstat_gen == ctx->store_gen && !A.cmax.empty()) return TOPK_OK;
    const uint32_t d = ctx->d;
    if (!A.cstat) TK_HIP(hipMalloc(&A.cstat, sizeof(uint32_t) * 2 * TOPK_MAX_DIMS));
    std::vector<uint32_t> init(2 * d);
    for (uint32_t m = 0; m < d; m++) {
        init[m] = okey(-INFINITY);
        init[d + m] = okey(INFINITY);
    }
    hipStream_t s = ctx->stream;
    TK_HIP(hipMemcpyAsync(A.cstat, init.data(), sizeof(uint32_t) * 2 * d, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(colstats_kernel, dim3(256, d), dim3(256), 0, s, ctx->cols, ctx->n, ctx->ld, A.cstat,
                       A.cstat + d);
    TK_HIP(hipGetLastError());
    TK_HIP(hipMemcpyAsync(init.data(), A.cstat, sizeof(uint32_t) * 2 * d, hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    A.cmax.resize(d);
    A.cmin.resize(d);
    for (uint32_t m = 0; m < d; m++) {
        A.cmax[m] = okey_inv(init[m]);
        A.cmin[m] = okey_inv(init[d + m]);
    }
    A.cstat_cols = ctx->cols;
    A.cstat_gen = ctx->store_gen;
    return TOPK_OK;
}

}  // namespace topk

using namespace topk;

extern "C" {

int topk_radix_select_u32(topk_ctx *ctx, const uint32_t *d_keys, uint64_t n, uint64_t r, uint32_t *out) {
    if (!ctx || !d_keys || !out || n == 0 || r >= n || n > 0xFFFFFFFFull) return TOPK_ERR_INVALID;
    AuxState &A = aux(ctx);
    int rc = ensure_dsel(A);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    dsel_enqueue(s, U32Src{d_keys}, nullptr, n, r, 0, A.dsel, ((uintptr_t)d_keys & 15u) == 0);
    timing_end(ctx, s, "radix_select", ev0);
    TK_HIP(hipGetLastError());
    DselState h;
    TK_HIP(hipMemcpyAsync(&h, A.dsel, 16, hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    *out = h.prefix;
    return TOPK_OK;
}

int topk_bta(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, uint64_t *out_ids,
             float *out_scores, float *out_threshold) {
    if (!ctx || k < 1) return TOPK_ERR_INVALID;
    if (k > TOPK_MAX_K || ctx->n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;  // run keys carry 32-bit ids
    QCols q;
    int rc = build_qcols(ctx, ctx->cols, ctx->ld, qq, weights, attr, q);
    if (rc) return rc;
    const double t0 = now_ms();
    AuxState &A = aux(ctx);
    // runs of 64 keys for k <= 64: the register-resident block kernel
    const uint32_t K2 = k <= 64 ? 64u : next_pow2(k);
    const uint32_t nblk = (uint32_t)((ctx->n + BTA_OBJS - 1) / BTA_OBJS);
    const uint64_t need = (uint64_t)nblk * K2;
    if (need > A.bta_cap) {
        hipFree(A.bta[0]);
        hipFree(A.bta[1]);
        A.bta[0] = A.bta[1] = nullptr;
        A.bta_cap = 0;
        TK_HIP(hipMalloc(&A.bta[0], sizeof(uint64_t) * need));
        TK_HIP(hipMalloc(&A.bta[1], sizeof(uint64_t) * need));
        A.bta_cap = need;
    }
    hipStream_t s = ctx->stream;
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    if (K2 == 64)
        hipLaunchKernelGGL(bta_block64_kernel, dim3(nblk), dim3(BTA_BS), 0, s, q, ctx->n, A.bta[0]);
    else
        hipLaunchKernelGGL(bta_block_kernel, dim3(nblk), dim3(BTA_BS), 0, s, q, ctx->n, K2, A.bta[0]);
    timing_end(ctx, s, "bta_block", ev0);
    uint32_t nruns = nblk, cur = 0;
    const uint32_t G = BTA_OBJS / K2;
    while (nruns > 1) {
        const uint32_t ng = (nruns + G - 1) / G;
        hipLaunchKernelGGL(bta_merge_kernel, dim3(ng), dim3(BTA_BS), 0, s, A.bta[cur], nruns, K2, A.bta[cur ^ 1]);
        cur ^= 1;
        nruns = ng;
    }
    const uint32_t c = (uint32_t)(ctx->n < k ? ctx->n : k);
    hipLaunchKernelGGL(emit_run_kernel, dim3(1), dim3(256), 0, s, A.bta[cur], c, k,
                       reinterpret_cast<uint64_t *>(ctx->d_res + ctx->res.ids_off),
                       reinterpret_cast<float *>(ctx->d_res + ctx->res.scores_off),
                       reinterpret_cast<float *>(ctx->d_res + ctx->res.thr_off),
                       reinterpret_cast<uint32_t *>(ctx->d_res + ctx->res.status_off));
    TK_HIP(hipGetLastError());
    rc = finish_single(ctx, s, k, out_ids, out_scores, out_threshold);
    if (rc) return rc;
    topk_stats &S = ctx->stats;
    S.queries = 1;
    S.fallback = 0;
    S.retried = 0;
    S.fte_fallback = 0;
    S.tuple_count = ctx->n;
    S.terms = ctx->n * qq;
    S.candidates = (uint64_t)nblk * K2;
    S.partitions_scored = nblk;
    S.partitions_total = nblk;
    S.ms_last = now_ms() - t0;
    return TOPK_OK;
}

int topk_gpa(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, uint64_t *out_ids,
             float *out_scores, float *out_threshold) {
    if (!ctx || k < 1) return TOPK_ERR_INVALID;
    if (k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    QCols q;
    int rc = build_qcols(ctx, ctx->cols, ctx->ld, qq, weights, attr, q);
    if (rc) return rc;
    if (ctx->n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;  // survivor ids are 32-bit
    const double t0 = now_ms();
    AuxState &A = aux(ctx);
    rc = ensure_dsel(A);
    if (!rc) rc = column_stats(ctx, A);
    if (rc) return rc;
    const uint64_t n = ctx->n;
    if (n > A.gcap) {
        for (int b = 0; b < 2; b++) {
            hipFree(A.gids[b]);
            hipFree(A.gs[b]);
            A.gids[b] = nullptr;
            A.gs[b] = nullptr;
        }
        A.gcap = 0;
        for (int b = 0; b < 2; b++) {
            TK_HIP(hipMalloc(&A.gids[b], sizeof(uint32_t) * n));
            TK_HIP(hipMalloc(&A.gs[b], sizeof(float) * n));
        }
        A.gcap = n;
    }
    if (!A.gcnt) TK_HIP(hipMalloc(&A.gcnt, sizeof(uint32_t) * 2));
    if (!A.gthr) TK_HIP(hipMalloc(&A.gthr, sizeof(float)));
    if (!A.gbcount) TK_HIP(hipMalloc(&A.gbcount, sizeof(uint32_t) * GPA_MAX_GRID));
    // remaining-term bounds after each term, FP64
    const uint32_t m = q.m;
    std::vector<GpaStep> steps(m);
    for (uint32_t j = 0; j < m; j++) {
        GpaStep g{0, 0, 0, j + 1 == m};
        for (uint32_t t = j + 1; t < m; t++) {
            const double w = (double)weights[attr[t]];
            const double hi = w * (double)A.cmax[attr[t]], lo = w * (double)A.cmin[attr[t]];
            g.hi += hi > lo ? hi : lo;
            g.lo += hi > lo ? lo : hi;
            g.r += fabs(w) * fmax(fabs((double)A.cmax[attr[t]]), fabs((double)A.cmin[attr[t]]));
        }
        steps[j] = g;
    }
    const uint32_t kreq = (uint32_t)(n < k ? n : k);
    hipStream_t s = ctx->stream;
    uint32_t g = (uint32_t)((n + GPA_BS - 1) / GPA_BS);
    if (g > GPA_MAX_GRID) g = GPA_MAX_GRID;
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    hipLaunchKernelGGL(gpa_first_kernel, dim3(g), dim3(GPA_BS), 0, s, q.col[0], q.w[0], n, A.gids[0], A.gs[0],
                       A.gcnt);
    int cur = 0;
    for (uint32_t j = 0; j < m; j++) {
        dsel_enqueue(s, ScoreSrc{A.gs[cur]}, A.gcnt + cur, n, 0, kreq, A.dsel, true);  // hipMalloc'd: aligned
        hipLaunchKernelGGL(gpa_thr_kernel, dim3(1), dim3(1), 0, s, A.dsel, steps[j], A.gthr);
        const bool more = j + 1 < m;
        hipLaunchKernelGGL(gpa_count_kernel, dim3(g), dim3(GPA_BS), 0, s, A.gs[cur], A.gcnt + cur, steps[j], A.gthr,
                           A.gbcount);
        hipLaunchKernelGGL(gpa_prune_kernel, dim3(g), dim3(GPA_BS), 0, s, A.gids[cur], A.gs[cur], A.gcnt + cur,
                           steps[j], A.gthr, A.gbcount, more ? q.col[j + 1] : (const float *)nullptr,
                           more ? q.w[j + 1] : 0.0f, A.gids[cur ^ 1], A.gs[cur ^ 1], A.gcnt + (cur ^ 1));
        cur ^= 1;
    }
    const FinalOut fo{reinterpret_cast<uint64_t *>(ctx->d_res + ctx->res.ids_off),
                      reinterpret_cast<float *>(ctx->d_res + ctx->res.scores_off),
                      reinterpret_cast<float *>(ctx->d_res + ctx->res.thr_off),
                      reinterpret_cast<uint32_t *>(ctx->d_res + ctx->res.status_off)};
    hipLaunchKernelGGL(gpa_final_kernel, dim3(1), dim3(512), 0, s, A.gs[cur], A.gids[cur], A.gcnt + cur, kreq, fo);
    timing_end(ctx, s, "gpa", ev0);
    TK_HIP(hipGetLastError());
    uint32_t last = 0;
    TK_HIP(hipMemcpyAsync(&last, A.gcnt + cur, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    rc = finish_single(ctx, s, k, out_ids, out_scores, out_threshold);
    if (rc) return rc;
    if (k > kreq) {
        for (uint32_t i = kreq; i < k; i++) {
            if (out_ids) out_ids[i] = ~0ull;
            if (out_scores) out_scores[i] = -INFINITY;
        }
    }
    topk_stats &S = ctx->stats;
    S.queries = 1;
    S.fallback = 0;
    S.retried = 0;
    S.fte_fallback = 0;
    S.candidates = last;
    S.tuple_count = n;
    S.terms = 0;  // attributes are scored progressively over shrinking sets: not counted
    S.partitions_scored = 0;
    S.partitions_total = 0;
    S.ms_last = now_ms() - t0;
    return TOPK_OK;
}

}  // extern "C"
