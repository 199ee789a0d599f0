// PTA: partitioned threshold top-k on gfx950.
//
// Reference: cpu/PTA.h:88-267 (partition build: polar angles, sorted lists,
// per-block threshold vectors) and cpu/PTA.h:357-422, 569-671 (scan with a
// per-block bound, stop once the heap top beats it).  This file keeps the
// idea -- group similar objects, bound each group, skip groups that cannot
// reach the running k-th score -- with its own partitioning and bound:
//
//  * layout (topk_init with TOPK_INIT_PTA): a k-d tree whose split planes
//    come from a deterministic sample (median splits on the axis of largest
//    spread, built on the host); every object is routed to its leaf on the
//    GPU, objects are counting-sorted by leaf, then ordered inside each
//    partition by attribute sum (largest first, ties by id); the columns are
//    gathered into partition order and each partition keeps per-attribute
//    max/min, the max of its objects' attribute sums and the sum of its
//    minima, and each 256-object chunk the same statistics over itself and
//    every later chunk of its partition (suffix rows).
//  * bound: the largest weighted sum any point of the partition's box can
//    reach under the attribute-sum limit (a fractional-knapsack LP, solved
//    greedily in FP64), plus a margin covering binary32 rounding of a score,
//    rounded up to binary32.  It is never below an object's score.
//  * query: bound keys for all partitions; the exact k-th key among the best
//    partitions' objects (seed) is a proven lower bound tau of the global
//    k-th key; the task pass lists the partitions whose bound reaches tau's
//    score (skip iff bound < tau strictly), each up to its first suffix row
//    whose bound falls below tau, as chunk ranges; the scoring pass spreads
//    those chunks evenly over every wave of the GPU and keeps keys >= tau;
//    the FTE final select writes the exact top k straight into pinned host
//    memory.  Results equal topk_fte's.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

#include "bound.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace topk {

constexpr int PTA_BS = 256;
constexpr uint32_t PTA_CUT_GROUPS = 4096;  // cut workgroups at most (the scoring pass scans their totals in LDS)
constexpr uint32_t PTA_SORT_MAX = 8192;      // partitions up to this size are sorted in LDS (larger: scatter order)
constexpr uint32_t PTA_SEED_OBJS = 1024;     // leading objects (largest attribute sums) of a seed partition
constexpr uint32_t PTA_MIN_SEED = 16;        // seed partitions at least
constexpr uint32_t PTA_SEED_RANK = 1024;     // block maxima ranked in LDS by the seed kernel
constexpr uint32_t PTA_MAX_SEED = 64;
constexpr int PTA_SEED_REG = 16;             // seed winners per thread the merge reads at once
constexpr uint32_t PTA_SEED_LDS = PTA_SEED_REG * PTA_BS;  // seed winners the merge ranks in LDS

struct KdNode {
    float split;
    uint32_t axis;
    int32_t left;   // >= 0: internal node, < 0: leaf -(leaf + 1)
    int32_t right;
};

// ---------------------------------------------------------------------------
// layout build kernels
// ---------------------------------------------------------------------------
__global__ void pta_sample_kernel(const float *__restrict__ cols, uint64_t ld, uint32_t d, uint64_t n, uint64_t s,
                                  float *__restrict__ rows) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= s) return;
    const uint64_t i = j * n / s;
    for (uint32_t m = 0; m < d; m++) rows[j * d + m] = cols[(uint64_t)m * ld + i];
}

__global__ void pta_assign_kernel(const float *__restrict__ cols, uint64_t ld, uint64_t n,
                                  const KdNode *__restrict__ nodes, int32_t root, uint32_t *__restrict__ leaf,
                                  uint32_t *__restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t c = root;
    while (c >= 0) {
        const KdNode nd = nodes[c];
        c = cols[(uint64_t)nd.axis * ld + i] < nd.split ? nd.left : nd.right;
    }
    const uint32_t l = (uint32_t)(-c - 1);
    leaf[i] = l;
    atomicAdd(&cnt[l], 1u);
}

__global__ void pta_scatter_kernel(uint64_t n, const uint32_t *__restrict__ leaf, const uint32_t *__restrict__ poff,
                                   uint32_t *__restrict__ cursor, uint32_t *__restrict__ perm) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t l = leaf[i];
    perm[poff[l] + atomicAdd(&cursor[l], 1u)] = (uint32_t)i;
}

// Sorts each partition's objects by attribute sum, largest first (ties by
// lower id): consecutive positions then hold objects of similar sum, so a
// chunk's statistics bound its objects more tightly than the partition's.
// The sum only orders; any rounding is fine.
__global__ __launch_bounds__(1024) void pta_sort_kernel(const float *__restrict__ cols, uint64_t ld, uint32_t d,
                                                        const uint32_t *__restrict__ poff,
                                                        const uint32_t *__restrict__ psz,
                                                        uint32_t *__restrict__ perm) {
    __shared__ uint64_t a[PTA_SORT_MAX];
    const uint32_t p = blockIdx.x, s = psz[p];
    if (s < 2 || s > PTA_SORT_MAX) return;
    uint32_t *seg = perm + poff[p];
    uint32_t n2 = 1;
    while (n2 < s) n2 <<= 1;
    for (uint32_t i = threadIdx.x; i < n2; i += 1024) {
        uint64_t key = 0;
        if (i < s) {
            const uint32_t id = seg[i];
            float sum = 0.0f;
            for (uint32_t m = 0; m < d; m++) sum += cols[(uint64_t)m * ld + id];
            key = rkey(sum, id);
        }
        a[i] = key;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= n2; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n2 / 2; t += 1024) {
                const uint32_t lo = 2 * stride * (t / stride) + (t & (stride - 1)), hi = lo + stride;
                const bool desc = (lo & size) == 0;
                const uint64_t x = a[lo], y = a[hi];
                if ((x < y) == desc) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < s; i += 1024) seg[i] = rkey_id(a[i]);
}

__global__ void pta_gather_kernel(const float *__restrict__ cols, uint64_t ld, uint32_t d, uint64_t npos,
                                  const uint32_t *__restrict__ perm, float *__restrict__ pcols, uint64_t ldp,
                                  uint32_t *__restrict__ pid) {
    const uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= npos) return;
    const uint32_t id = perm[pos];
    pid[pos] = id;
    for (uint32_t m = 0; m < d; m++) pcols[(uint64_t)m * ldp + pos] = id != 0xFFFFFFFFu ? cols[(uint64_t)m * ld + id] : 0.0f;
}

// One block per partition: per-attribute max/min, the max attribute sum
// (FP64, rounded up) and the sum of the minima (FP64, rounded down).
__global__ __launch_bounds__(PTA_BS) void pta_bounds_kernel(const float *__restrict__ pcols, uint64_t ldp, uint32_t d,
                                                            const uint32_t *__restrict__ poff,
                                                            const uint32_t *__restrict__ psz,
                                                            float *__restrict__ pmax, float *__restrict__ pmin,
                                                            float *__restrict__ psum, float *__restrict__ pminsum) {
    __shared__ float smax[PTA_BS / 64], smin[PTA_BS / 64];
    __shared__ double ssum[PTA_BS / 64];
    __shared__ float colmin[TOPK_MAX_DIMS];
    const uint32_t p = blockIdx.x, start = poff[p], s = psz[p];
    const uint32_t w = threadIdx.x / 64;
    for (uint32_t m = 0; m < d; m++) {
        float mx = -INFINITY, mn = INFINITY;
        for (uint32_t i = threadIdx.x; i < s; i += PTA_BS) {
            const float v = pcols[(uint64_t)m * ldp + start + i];
            mx = v > mx ? v : mx;
            mn = v < mn ? v : mn;
        }
        for (int o = 32; o > 0; o >>= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            mn = fminf(mn, __shfl_xor(mn, o, 64));
        }
        if (lane_id() == 0) { smax[w] = mx; smin[w] = mn; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = 1; j < PTA_BS / 64; j++) { mx = fmaxf(mx, smax[j]); mn = fminf(mn, smin[j]); }
            pmax[(uint64_t)p * d + m] = s ? mx : 0.0f;
            pmin[(uint64_t)p * d + m] = s ? mn : 0.0f;
            colmin[m] = s ? mn : 0.0f;
        }
        __syncthreads();
    }
    double best = -INFINITY;
    for (uint32_t i = threadIdx.x; i < s; i += PTA_BS) {
        double t = 0;
        for (uint32_t m = 0; m < d; m++) t += (double)pcols[(uint64_t)m * ldp + start + i];
        best = t > best ? t : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(best, o, 64);
        best = t > best ? t : best;
    }
    if (lane_id() == 0) ssum[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < PTA_BS / 64; j++) best = ssum[j] > best ? ssum[j] : best;
        double ms = 0;
        for (uint32_t m = 0; m < d; m++) ms += (double)colmin[m];
        psum[p] = s ? f32_round_up(best) : -INFINITY;
        pminsum[p] = s ? f32_round_down(ms) : 0.0f;
    }
}

// One wave per chunk of PTA_CHUNK positions: element e of chunk c (at
// cstat[e * nchunks + c]) holds the chunk's per-attribute maxima (e in
// [0, d)) and minima ([d, 2d)), the largest attribute sum of its objects
// (FP64, rounded up) at 2d and the sum of its minima (FP64, rounded down)
// at 2d + 1.  A chunk without objects has sum -inf.  cminid[c] gets the
// chunk's smallest object id (0xFFFFFFFF: none).
__global__ __launch_bounds__(256) void pta_chunk_kernel(const float *__restrict__ pcols, uint64_t ldp, uint32_t d,
                                                        const uint32_t *__restrict__ pid, uint64_t nchunks,
                                                        float *__restrict__ cstat, uint32_t *__restrict__ cminid) {
    const uint64_t c = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (c >= nchunks) return;
    const uint32_t lane = lane_id();
    constexpr uint32_t PER = PTA_CHUNK / 64;
    const uint64_t pos0 = c * PTA_CHUNK + lane * PER;
    bool ok[PER];
    bool any = false;
    uint32_t idmin = 0xFFFFFFFFu;  // padding positions hold id 0xFFFFFFFF
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) {
        const uint32_t id = pid[pos0 + j];
        ok[j] = id != 0xFFFFFFFFu;
        any |= ok[j];
        idmin = id < idmin ? id : idmin;
    }
    const bool nonempty = __ballot(any) != 0;
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t t = (uint32_t)__shfl_xor((int)idmin, o, 64);
        idmin = t < idmin ? t : idmin;
    }
    if (lane == 0) cminid[c] = idmin;
    float *row = cstat + c;  // element e at row[e * nchunks]
    double best = -INFINITY, minsum = 0.0;
    double sums[PER];
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) sums[j] = 0.0;
    for (uint32_t m = 0; m < d; m++) {
        float mx = -INFINITY, mn = INFINITY;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            const float v = pcols[(uint64_t)m * ldp + pos0 + j];
            if (ok[j]) {
                mx = v > mx ? v : mx;
                mn = v < mn ? v : mn;
                sums[j] += (double)v;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            mn = fminf(mn, __shfl_xor(mn, o, 64));
        }
        minsum += (double)mn;
        if (lane == 0) {
            row[m * nchunks] = nonempty ? mx : 0.0f;
            row[(d + m) * nchunks] = nonempty ? mn : 0.0f;
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) best = ok[j] && sums[j] > best ? sums[j] : best;
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(best, o, 64);
        best = t > best ? t : best;
    }
    if (lane == 0) {
        row[2 * d * nchunks] = nonempty ? f32_round_up(best) : -INFINITY;
        row[(2 * d + 1) * nchunks] = nonempty ? f32_round_down(minsum) : 0.0f;
    }
}

// One thread per partition turns its chunk rows into suffix rows: row c
// then bounds chunks c, c + 1, ... of the partition (per-attribute maxima and
// minima, the largest attribute sum, and the sum of the union's minima in
// FP64 rounded down), so a partition's row bounds never increase with c.
__global__ __launch_bounds__(256) void pta_chunk_suffix_kernel(uint32_t L, uint32_t d,
                                                               const uint32_t *__restrict__ poff,
                                                               const uint32_t *__restrict__ psz, uint64_t nchunks,
                                                               float *__restrict__ cstat) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= L || psz[p] == 0) return;
    const uint32_t c0 = poff[p] / PTA_CHUNK, n = (psz[p] + PTA_CHUNK - 1) / PTA_CHUNK;
    const uint64_t S = nchunks;  // element stride
    for (uint32_t i = n - 1; i > 0; i--) {
        const float *nx = cstat + (c0 + i);
        float *row = cstat + (c0 + i - 1);
        double ms = 0.0;
        for (uint32_t m = 0; m < d; m++) {
            row[m * S] = fmaxf(row[m * S], nx[m * S]);
            row[(d + m) * S] = fminf(row[(d + m) * S], nx[(d + m) * S]);
            ms += (double)row[(d + m) * S];
        }
        row[2 * d * S] = fmaxf(row[2 * d * S], nx[2 * d * S]);
        row[(2 * d + 1) * S] = f32_round_down(ms);
    }
}

// ---------------------------------------------------------------------------
// host: k-d tree on the sample
// ---------------------------------------------------------------------------
namespace {

struct KdBuilder {
    const float *rows;
    uint32_t d;
    std::vector<uint32_t> idx;
    std::vector<KdNode> nodes;  // internal node for the split between leaves m-1 and m is nodes[m-1]

    // Builds the subtree holding leaves [leaf_lo, leaf_lo + nl) from sample
    // rows idx[lo, hi); returns its node code.
    int32_t build(size_t lo, size_t hi, uint32_t leaf_lo, uint32_t nl, int depth) {
        if (nl == 1) return -(int32_t)leaf_lo - 1;
        const uint32_t nleft = nl / 2;
        const size_t mid = lo + (hi - lo) * nleft / nl;
        uint32_t axis = 0;
        float split = INFINITY;
        if (hi > lo) {
            float best = -1.0f;
            for (uint32_t m = 0; m < d; m++) {
                float mn = INFINITY, mx = -INFINITY;
                for (size_t j = lo; j < hi; j++) {
                    const float v = rows[(size_t)idx[j] * d + m];
                    mn = v < mn ? v : mn;
                    mx = v > mx ? v : mx;
                }
                if (mx - mn > best) { best = mx - mn; axis = m; }
            }
            if (mid < hi) {
                auto less = [&](uint32_t a, uint32_t b) {
                    const float va = rows[(size_t)a * d + axis], vb = rows[(size_t)b * d + axis];
                    return va < vb || (va == vb && a < b);
                };
                std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, less);
                split = rows[(size_t)idx[mid] * d + axis];
            }
        }
        const int32_t me = (int32_t)(leaf_lo + nleft - 1);
        int32_t l, r;
        if (depth < 4 && hi - lo > 65536) {
            std::thread th([&] { l = build(lo, mid, leaf_lo, nleft, depth + 1); });
            r = build(mid, hi, leaf_lo + nleft, nl - nleft, depth + 1);
            th.join();
        } else {
            l = build(lo, mid, leaf_lo, nleft, depth + 1);
            r = build(mid, hi, leaf_lo + nleft, nl - nleft, depth + 1);
        }
        nodes[me] = KdNode{split, axis, l, r};
        return me;
    }
};

}  // namespace

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static float round_down_f32(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

// Aggregates the per-partition bounds bottom-up into the bound tree.
static int build_bound_tree(PtaLayout &P, uint32_t d, hipStream_t s) {
    const uint32_t L = P.nparts;
    uint32_t p2 = 1;
    while (p2 < L) p2 <<= 1;
    const size_t nn = 2 * (size_t)p2;
    std::vector<float> mx(nn * d, -INFINITY), mn(nn * d, INFINITY), sm(nn, -INFINITY), ms(nn, 0.0f);
    std::vector<uint32_t> sz(L), off(L + 1), cmin(P.nchunks), idm(nn, 0xFFFFFFFFu);
    TK_HIP(hipMemcpyAsync(mx.data() + (size_t)p2 * d, P.pmax, sizeof(float) * L * d, hipMemcpyDeviceToHost, s));
    TK_HIP(hipMemcpyAsync(mn.data() + (size_t)p2 * d, P.pmin, sizeof(float) * L * d, hipMemcpyDeviceToHost, s));
    TK_HIP(hipMemcpyAsync(sm.data() + p2, P.psum, sizeof(float) * L, hipMemcpyDeviceToHost, s));
    TK_HIP(hipMemcpyAsync(ms.data() + p2, P.pminsum, sizeof(float) * L, hipMemcpyDeviceToHost, s));
    TK_HIP(hipMemcpyAsync(sz.data(), P.psz, sizeof(uint32_t) * L, hipMemcpyDeviceToHost, s));
    TK_HIP(hipMemcpyAsync(off.data(), P.poff, sizeof(uint32_t) * (L + 1), hipMemcpyDeviceToHost, s));
    if (P.nchunks)
        TK_HIP(hipMemcpyAsync(cmin.data(), P.cminid, sizeof(uint32_t) * P.nchunks, hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    // smallest id per partition (its chunks), then per node
    for (uint32_t p = 0; p < L; p++) {
        uint32_t m = 0xFFFFFFFFu;
        if (sz[p])
            for (uint64_t c = off[p] / PTA_CHUNK; c < ((uint64_t)off[p] + sz[p] + PTA_CHUNK - 1) / PTA_CHUNK; c++)
                m = std::min(m, cmin[c]);
        idm[p2 + p] = m;
    }
    for (size_t v = p2 - 1; v >= 1; v--) idm[v] = std::min(idm[2 * v], idm[2 * v + 1]);
    for (uint32_t p = 0; p < L; p++)
        if (!sz[p]) {  // empty partition: identity box
            for (uint32_t m = 0; m < d; m++) {
                mx[(size_t)(p2 + p) * d + m] = -INFINITY;
                mn[(size_t)(p2 + p) * d + m] = INFINITY;
            }
            sm[p2 + p] = -INFINITY;
        }
    for (size_t v = p2 - 1; v >= 1; v--) {
        const size_t a = 2 * v, b = 2 * v + 1;
        bool any = false;
        for (uint32_t m = 0; m < d; m++) {
            mx[v * d + m] = std::max(mx[a * d + m], mx[b * d + m]);
            mn[v * d + m] = std::min(mn[a * d + m], mn[b * d + m]);
        }
        sm[v] = std::max(sm[a], sm[b]);
        any = sm[v] > -INFINITY;
        double t = 0;
        for (uint32_t m = 0; m < d; m++) t += any ? (double)mn[v * d + m] : 0.0;
        ms[v] = any ? round_down_f32(t) : 0.0f;
    }
    P.p2 = p2;
    // attribute-major copies for the device: attribute m of node v at [m * nn + v]
    std::vector<float> mxt(nn * d), mnt(nn * d);
    for (size_t v = 0; v < nn; v++)
        for (uint32_t m = 0; m < d; m++) {
            mxt[m * nn + v] = mx[v * d + m];
            mnt[m * nn + v] = mn[v * d + m];
        }
    TK_HIP(hipMalloc(&P.tmax, sizeof(float) * nn * d));
    TK_HIP(hipMalloc(&P.tmin, sizeof(float) * nn * d));
    TK_HIP(hipMalloc(&P.tsum, sizeof(float) * nn));
    TK_HIP(hipMalloc(&P.tminsum, sizeof(float) * nn));
    TK_HIP(hipMalloc(&P.tminid, sizeof(uint32_t) * nn));
    TK_HIP(hipMemcpyAsync(P.tminid, idm.data(), sizeof(uint32_t) * nn, hipMemcpyHostToDevice, s));
    TK_HIP(hipMemcpyAsync(P.tmax, mxt.data(), sizeof(float) * nn * d, hipMemcpyHostToDevice, s));
    TK_HIP(hipMemcpyAsync(P.tmin, mnt.data(), sizeof(float) * nn * d, hipMemcpyHostToDevice, s));
    TK_HIP(hipMemcpyAsync(P.tsum, sm.data(), sizeof(float) * nn, hipMemcpyHostToDevice, s));
    TK_HIP(hipMemcpyAsync(P.tminsum, ms.data(), sizeof(float) * nn, hipMemcpyHostToDevice, s));
    TK_HIP(hipStreamSynchronize(s));
    return TOPK_OK;
}

static void pta_free(PtaLayout &P) {
    hipFree(P.pcols);
    hipFree(P.pid);
    hipFree(P.poff);
    hipFree(P.psz);
    hipFree(P.pmax);
    hipFree(P.pmin);
    hipFree(P.psum);
    hipFree(P.pminsum);
    hipFree(P.pbound);
    hipFree(P.bmax);
    hipFree(P.pcut);
    hipFree(P.wchunks);
    hipFree(P.wstat);
    if (P.h_report) hipHostFree(P.h_report);
    hipFree(P.tmax);
    hipFree(P.tmin);
    hipFree(P.tsum);
    hipFree(P.tminsum);
    hipFree(P.cstat);
    hipFree(P.cminid);
    hipFree(P.tminid);
    hipFree(P.xbits);
    const uint32_t psize = P.psize;
    const uint64_t gen = P.generation;
    P = PtaLayout();
    P.psize = psize;
    P.generation = gen + 1;
}

// One wave per (row, word): lane i reads object 64 w + i of the attribute
// and the ballot of "equals the store's largest / smallest value" is the word.
__global__ __launch_bounds__(256) void extreme_bits_kernel(const float *__restrict__ cols, uint64_t ld, uint64_t n,
                                                           uint32_t d, const float *__restrict__ rmax,
                                                           const float *__restrict__ rmin, uint64_t rstride,
                                                           uint64_t nw, uint64_t *__restrict__ bits) {
    const uint32_t lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    for (uint64_t v = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64; v < (uint64_t)d * nw; v += nwaves) {
        const uint32_t a = (uint32_t)(v / nw);
        const uint64_t w = v - (uint64_t)a * nw, id = 64 * w + lane;
        const float x = id < n ? cols[(uint64_t)a * ld + id] : 0.0f;
        const uint64_t hi = __ballot(id < n && x == rmax[a * rstride]);
        const uint64_t lo = __ballot(id < n && x == rmin[a * rstride]);
        if (lane == 0) {
            bits[(uint64_t)a * nw + w] = hi;
            bits[(uint64_t)(d + a) * nw + w] = lo;
        }
    }
}

int pta_extreme_bits(topk_ctx *ctx) {
    PtaLayout &P = ctx->pta;
    if (!P.built) return TOPK_ERR_STATE;
    if (P.xbits) return TOPK_OK;
    const uint64_t nw = ((ctx->n + 63) / 64 + 15) & ~15ull;  // padding words hold ids >= n: zero
    TK_HIP(hipMalloc(&P.xbits, sizeof(uint64_t) * 2 * ctx->d * nw));
    P.nw = nw;
    const uint64_t waves = (uint64_t)ctx->d * nw;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((waves + 3) / 4, 65536);
    hipLaunchKernelGGL(extreme_bits_kernel, dim3(grid), dim3(256), 0, ctx->stream, ctx->cols, ctx->ld, ctx->n, ctx->d,
                       P.tmax + 1, P.tmin + 1, 2 * (uint64_t)P.p2, nw, P.xbits);
    TK_HIP(hipGetLastError());
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TOPK_OK;
}

void pta_release(topk_ctx *ctx) {
    batch_quiesce(ctx);  // batches in flight read the layout
    pta_free(ctx->pta);
}

int pta_build(topk_ctx *ctx) {
    const double t0 = now_ms();
    PtaLayout &P = ctx->pta;
    pta_release(ctx);
    const uint64_t n = ctx->n;
    const uint32_t d = ctx->d;
    hipStream_t s = ctx->stream;
    if (n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;  // positions and partition ids are 32-bit
    const uint64_t want = (n + P.psize - 1) / P.psize;
    const uint32_t L = (uint32_t)(want < PTA_MAX_PARTS ? want : PTA_MAX_PARTS);

    // 1. deterministic sample -> host k-d tree with L leaves
    uint64_t ns = (uint64_t)L * 64;
    if (ns < (1u << 20)) ns = 1u << 20;
    if (ns > n) ns = n;
    std::vector<float> rows((size_t)ns * d);
    {
        float *d_rows = nullptr;
        TK_HIP(hipMalloc(&d_rows, sizeof(float) * ns * d));
        hipLaunchKernelGGL(pta_sample_kernel, dim3((uint32_t)((ns + 255) / 256)), dim3(256), 0, s, ctx->cols,
                           ctx->ld, d, n, ns, d_rows);
        hipError_t e = hipMemcpyAsync(rows.data(), d_rows, sizeof(float) * ns * d, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        hipFree(d_rows);
        TK_HIP(e);
    }
    KdBuilder kb{rows.data(), d, std::vector<uint32_t>(ns), std::vector<KdNode>(L > 1 ? L - 1 : 1)};
    std::iota(kb.idx.begin(), kb.idx.end(), 0u);
    const int32_t root = kb.build(0, ns, 0, L, 0);

    // 2. route every object to its leaf and count
    KdNode *d_nodes = nullptr;
    uint32_t *leaf = nullptr, *cnt = nullptr, *cursor = nullptr, *perm = nullptr;
    int rc = TOPK_OK;
    auto cleanup = [&] {
        hipFree(d_nodes);
        hipFree(leaf);
        hipFree(cnt);
        hipFree(cursor);
    };
#define TK_TRY(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) { rc = hip_status(e_); goto fail; } \
    } while (0)
    {
        TK_TRY(hipMalloc(&d_nodes, sizeof(KdNode) * kb.nodes.size()));
        TK_TRY(hipMemcpyAsync(d_nodes, kb.nodes.data(), sizeof(KdNode) * kb.nodes.size(), hipMemcpyHostToDevice, s));
        TK_TRY(hipMalloc(&leaf, sizeof(uint32_t) * n));
        TK_TRY(hipMalloc(&cnt, sizeof(uint32_t) * L));
        TK_TRY(hipMalloc(&cursor, sizeof(uint32_t) * L));
        TK_TRY(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * L, s));
        TK_TRY(hipMemsetAsync(cursor, 0, sizeof(uint32_t) * L, s));
        hipLaunchKernelGGL(pta_assign_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, ctx->cols, ctx->ld, n,
                           d_nodes, root, leaf, cnt);
        std::vector<uint32_t> h_cnt(L), h_off(L + 1);
        TK_TRY(hipMemcpyAsync(h_cnt.data(), cnt, sizeof(uint32_t) * L, hipMemcpyDeviceToHost, s));
        TK_TRY(hipStreamSynchronize(s));
        uint64_t acc = 0;
        for (uint32_t p = 0; p < L; p++) {
            P.nonempty += h_cnt[p] ? 1u : 0u;
            P.maxsz = h_cnt[p] > P.maxsz ? h_cnt[p] : P.maxsz;
            h_off[p] = (uint32_t)acc;
            acc += ((uint64_t)h_cnt[p] + PART_ALIGN - 1) / PART_ALIGN * PART_ALIGN;
        }
        h_off[L] = (uint32_t)acc;
        if (acc >= 0xFFFFFFFFull) { rc = TOPK_ERR_UNSUPPORTED; goto fail; }
        P.nparts = L;
        P.npos = acc;
        P.ld = (acc + 8191) / 8192 * 8192;
        TK_TRY(hipMalloc(&P.poff, sizeof(uint32_t) * (L + 1)));
        TK_TRY(hipMalloc(&P.psz, sizeof(uint32_t) * L));
        TK_TRY(hipMemcpyAsync(P.poff, h_off.data(), sizeof(uint32_t) * (L + 1), hipMemcpyHostToDevice, s));
        TK_TRY(hipMemcpyAsync(P.psz, h_cnt.data(), sizeof(uint32_t) * L, hipMemcpyHostToDevice, s));
        TK_TRY(hipStreamSynchronize(s));  // host vectors go out of scope

        // 3. counting sort by leaf, then by attribute sum inside each partition
        TK_TRY(hipMalloc(&perm, sizeof(uint32_t) * P.ld));
        TK_TRY(hipMemsetAsync(perm, 0xFF, sizeof(uint32_t) * P.ld, s));
        hipLaunchKernelGGL(pta_scatter_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, n, leaf, P.poff,
                           cursor, perm);
        hipLaunchKernelGGL(pta_sort_kernel, dim3(L), dim3(1024), 0, s, ctx->cols, ctx->ld, d, P.poff, P.psz, perm);

        // 4. partition-order columns and bounds
        const size_t pfl = P.ld * d + 8192;
        TK_TRY(hipMalloc(&P.pcols, sizeof(float) * pfl));
        TK_TRY(hipMemsetAsync(P.pcols, 0, sizeof(float) * pfl, s));
        P.pid = perm;
        perm = nullptr;
        hipLaunchKernelGGL(pta_gather_kernel, dim3((uint32_t)((P.ld + 255) / 256)), dim3(256), 0, s, ctx->cols,
                           ctx->ld, d, P.ld, P.pid, P.pcols, P.ld, P.pid);
        TK_TRY(hipMalloc(&P.pmax, sizeof(float) * L * d));
        TK_TRY(hipMalloc(&P.pmin, sizeof(float) * L * d));
        TK_TRY(hipMalloc(&P.psum, sizeof(float) * L));
        TK_TRY(hipMalloc(&P.pminsum, sizeof(float) * L));
        hipLaunchKernelGGL(pta_bounds_kernel, dim3(L), dim3(PTA_BS), 0, s, P.pcols, P.ld, d, P.poff, P.psz, P.pmax,
                           P.pmin, P.psum, P.pminsum);
        P.nchunks = P.npos / PTA_CHUNK;
        P.crow = 2 * d + 2;
        TK_TRY(hipMalloc(&P.cstat, sizeof(float) * (P.nchunks ? P.nchunks : 1) * P.crow));
        TK_TRY(hipMalloc(&P.cminid, sizeof(uint32_t) * (P.nchunks ? P.nchunks : 1)));
        if (P.nchunks)
            hipLaunchKernelGGL(pta_chunk_kernel, dim3((uint32_t)((P.nchunks + 3) / 4)), dim3(256), 0, s, P.pcols, P.ld,
                               d, P.pid, (uint64_t)P.nchunks, P.cstat, P.cminid);
        if (P.nchunks)
            hipLaunchKernelGGL(pta_chunk_suffix_kernel, dim3((L + 255) / 256), dim3(256), 0, s, L, d, P.poff, P.psz,
                               (uint64_t)P.nchunks, P.cstat);
        TK_TRY(hipMalloc(&P.pbound, sizeof(uint64_t) * L));
        TK_TRY(hipMalloc(&P.bmax, sizeof(uint64_t) * ((L + PTA_BS - 1) / PTA_BS)));
        TK_TRY(hipMalloc(&P.pcut, sizeof(uint32_t) * L));
        TK_TRY(hipMalloc(&P.wchunks, sizeof(uint32_t) * PTA_CUT_GROUPS));
        TK_TRY(hipMalloc(&P.wstat, sizeof(uint2) * PTA_CUT_GROUPS));
        TK_TRY(hipHostMalloc(&P.h_report, sizeof(uint32_t) * 4, hipHostMallocDefault));
        TK_TRY(hipDeviceGetAttribute(&P.cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        if (P.cus < 1) P.cus = 1;
        TK_TRY(hipGetLastError());
        TK_TRY(hipStreamSynchronize(s));
        rc = build_bound_tree(P, d, s);
        if (rc) goto fail;
    }
#undef TK_TRY
    cleanup();
    P.built = true;
    P.tt_build_ms = now_ms() - t0;
    return TOPK_OK;
fail:
    cleanup();
    hipFree(perm);
    pta_free(P);
    return rc;
}

// ---------------------------------------------------------------------------
// query kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PTA_BS) void pta_bound_kernel(BoundArgs b, uint32_t L,
                                                           const uint32_t *__restrict__ psz,
                                                           const float *__restrict__ pmax,
                                                           const float *__restrict__ pmin,
                                                           const float *__restrict__ psum,
                                                           const float *__restrict__ pminsum,
                                                           uint64_t *__restrict__ pbound, uint64_t *__restrict__ bmax) {
    __shared__ uint64_t wm[PTA_BS / 64];
    const uint32_t p = blockIdx.x * PTA_BS + threadIdx.x;
    uint64_t key = 0;
    if (p < L && psz[p]) {
        const float ub = lp_bound(b, pmax + (uint64_t)p * b.d, pmin + (uint64_t)p * b.d, psum[p], pminsum[p]);
        key = rkey(ub, p);
    }
    if (p < L) pbound[p] = key;
    const uint64_t m = wave_max_u64(key);
    if (lane_id() == 0) wm[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t r = wm[0];
        for (int j = 1; j < PTA_BS / 64; j++) r = wm[j] > r ? wm[j] : r;
        bmax[blockIdx.x] = r;
    }
}

// Seed winners blk[0..cnt) as a key source: KPT loads per thread issued
// together, the same trip count on every thread (ballots inside).
template <int KPT>
struct SeedKeys {
    const uint64_t *p;
    uint32_t cnt;
    template <class F>
    __device__ __forceinline__ void for_each(F f) const {
        uint64_t v[KPT];
#pragma unroll
        for (int j = 0; j < KPT; j++) {
            const uint32_t i = j * PTA_BS + threadIdx.x;
            v[j] = i < cnt ? p[i] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < KPT; j++) f(v[j]);
    }
};

// Block t scores the leading PTA_SEED_OBJS objects (largest attribute sums)
// of the partition whose bound key is the (t+1)-th largest of bmax[0..nb)
// (the partitions' own bound keys on small layouts, else the maxima of
// blocks of PTA_BS partitions) and keeps its top k; the last block to
// finish takes the k-th largest of all seed keys as tau.
template <int M>
__global__ __launch_bounds__(PTA_BS) void pta_seed_kernel(QCols q, const uint64_t *__restrict__ bmax, uint32_t nb,
                                                          const uint32_t *__restrict__ poff,
                                                          const uint32_t *__restrict__ psz,
                                                          const uint32_t *__restrict__ pid, uint32_t k,
                                                          uint64_t *__restrict__ blk, uint64_t *__restrict__ bthr,
                                                          FteState *__restrict__ st) {
    __shared__ SelectScratch ss;
    __shared__ uint64_t pick;
    __shared__ uint32_t is_last, nvalid;
    __shared__ uint64_t bm[PTA_SEED_RANK];
    __shared__ uint64_t cands[PTA_SEED_LDS];
    __shared__ uint64_t kth;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) pick = 0;
    if (nb <= PTA_SEED_RANK) {
        // the (t+1)-th largest block maximum: the nonzero key with exactly t
        // larger ones (keys carry partition ids, so nonzero keys are distinct)
        for (uint32_t i = tid; i < nb; i += PTA_BS) bm[i] = bmax[i];
        __syncthreads();
        for (uint32_t i = tid; i < nb; i += PTA_BS) {
            const uint64_t v = bm[i];
            uint32_t r = 0, unused = 0;
            rank2_lds(bm, nb, v, v, r, unused);
            if (v && r == blockIdx.x) pick = v;
        }
    } else if (tid < 64) {
        // (t+1)-th largest block maximum by wave-max rounds
        uint64_t best = 0;
        for (uint32_t r = 0; r <= blockIdx.x; r++) {
            uint64_t lm = 0;
            for (uint32_t i = lane_id(); i < nb; i += 64) {
                const uint64_t v = bmax[i];
                if (v > lm && (r == 0 || v < best)) lm = v;
            }
            best = wave_max_u64(lm);
        }
        if (tid == 0) pick = best;
    }
    __syncthreads();
    const uint64_t pk = pick;
    const uint32_t p = rkey_id(pk);
    const uint32_t start = pk ? poff[p] : 0u;
    const uint32_t size = pk ? (psz[p] < PTA_SEED_OBJS ? psz[p] : PTA_SEED_OBJS) : 0u;
    RegKeys<PTA_SEED_OBJS / PTA_BS> rk;
#pragma unroll
    for (int j = 0; j < (int)(PTA_SEED_OBJS / (4 * PTA_BS)); j++) {
        const uint32_t lo = j * 4 * PTA_BS + tid * 4;
        f4 sc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (lo < size) {
            const uint64_t v = (start + lo) / 4;
            if constexpr (M > 0) {
#pragma unroll
                for (int c = 0; c < M; c++) {
                    f4 pr = ld_stream(q.col[c] + v * 4) * q.w[c];
                    sc = sc + pr;
                }
            } else {
                for (uint32_t c = 0; c < q.m; c++) {
                    f4 pr = ld_stream(q.col[c] + v * 4) * q.w[c];
                    sc = sc + pr;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) rk.k[j * 4 + e] = lo + e < size ? rkey(sc[e], pid[start + lo + e]) : 0ull;
    }
    uint64_t thr = 1;
    if (size > k) thr = block_kth_largest<PTA_BS, RegKeys<PTA_SEED_OBJS / PTA_BS>, 11>(rk, size, k, ss);
    uint64_t *out = blk + (uint64_t)blockIdx.x * k;
    const uint32_t got = block_gather_ge<PTA_BS>(rk, thr, out, ss);
    for (uint32_t i = got + tid; i < k; i += PTA_BS) out[i] = 0;
    // this block's k-th key (k keys at or above it), or 0 with fewer than k
    if (tid == 0) bthr[blockIdx.x] = got >= k ? thr : 0ull;

    __threadfence();
    __syncthreads();
    if (tid == 0) is_last = (atomicAdd(&st->arrive, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    const uint32_t nk = gridDim.x * k;
    if (tid == 0) nvalid = 0;
    __syncthreads();
    uint64_t tau = 1;
    if (nk <= PTA_SEED_LDS) {
        // The union's k-th key is at least the largest block k-th key (that
        // block alone has k keys at or above it): only keys at or above it
        // compete.  They are gathered into LDS and the one with exactly k-1
        // larger keys is the k-th (keys are distinct).
        uint64_t floor_key = 0;
        for (uint32_t b = 0; b < gridDim.x; b++) floor_key = bthr[b] > floor_key ? bthr[b] : floor_key;
        const SeedKeys<PTA_SEED_REG> wk{blk, nk};
        const uint32_t nc = block_gather_ge<PTA_BS>(wk, floor_key ? floor_key : 1ull, cands, ss);
        if (nc > k && nc <= 2 * PTA_BS) {
            // two keys per thread, each ranked against all nc
            const uint64_t x = tid < nc ? cands[tid] : 0ull, y = tid + PTA_BS < nc ? cands[tid + PTA_BS] : 0ull;
            uint32_t rx = 0, ry = 0;
            rank2_lds(cands, nc, x, y, rx, ry);
            if (x && rx == k - 1) kth = x;
            if (y && ry == k - 1) kth = y;
            __syncthreads();
            tau = kth;
        } else if (nc > k) {
            tau = block_kth_largest<PTA_BS, SeedKeys<PTA_SEED_REG>, 11>(SeedKeys<PTA_SEED_REG>{cands, nc}, nc, k, ss);
        } else if (nc == k && floor_key) {
            tau = floor_key;  // exactly the largest block's k keys: its k-th is the union's
        }
    } else {
        uint32_t mine = 0;
        for (uint32_t i = tid; i < nk; i += PTA_BS) mine += blk[i] != 0;
        atomicAdd(&nvalid, mine);
        __syncthreads();
        if (nvalid > k) tau = block_kth_largest<PTA_BS>(MemKeys<PTA_BS>{blk, nk}, nvalid, k, ss);
    }
    if (tid == 0) {
        st->tau = tau;
        st->pre_count = 0;
        st->arrive = 0;
    }
}

// Cut per partition, as the number of its leading objects to score: 0 when
// its bound is below tau's score (skip iff bound < tau strictly), else the
// objects before its first chunk whose suffix-row bound falls below tau
// (every later object of the partition scores below tau too).  Workgroup
// g owns partitions [g * R, (g + 1) * R) (R a multiple of 16), four per
// wave per round, 16 lanes each, one chunk row per lane (a 4096-object
// partition in one round), and publishes its chunk total (wchunks[g]) and
// its partitions and objects to score (wstat[g]).  No atomics: the
// scoring pass scans the workgroup totals.
template <int NB>
__global__ __launch_bounds__(PTA_BS) void pta_cut_kernel(BoundArgs b, uint32_t L, uint32_t R,
                                                         const uint64_t *__restrict__ pbound,
                                                         const uint32_t *__restrict__ poff,
                                                         const uint32_t *__restrict__ psz,
                                                         const float *__restrict__ cstat, uint64_t nchunks,
                                                         const FteState *__restrict__ st, uint32_t *__restrict__ cut,
                                                         uint32_t *__restrict__ wchunks, uint2 *__restrict__ wstat) {
    __shared__ uint32_t wsum[PTA_BS / 64][3];
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64, g = lane / 16, gl = lane % 16;
    const uint32_t tau_hi = (uint32_t)(st->tau >> 32);
    const uint32_t d = b.d;
    uint32_t ch = 0, parts = 0, objs = 0;  // this lane's partitions (lanes gl == 0)
    for (uint32_t r = 0; r < R; r += 16) {
        const uint32_t p = blockIdx.x * R + r + wave * 4 + g;
        const uint64_t pb = p < L ? pbound[p] : 0ull;  // key 0: an empty partition
        const bool alive = pb != 0 && (uint32_t)(pb >> 32) >= tau_hi;
        const uint32_t sz = alive ? psz[p] : 0u;
        const uint32_t c0 = alive ? poff[p] / PTA_CHUNK : 0u;
        const uint32_t nc = (sz + PTA_CHUNK - 1) / PTA_CHUNK;
        uint32_t first = nc;  // this group's answer, settled once `done`
        bool done = !alive;
        for (uint32_t cb = 0; __ballot(!done); cb += 16) {
            const uint32_t c = cb + gl;
            bool fail = false;
            if (!done && c < nc) {
                const float *row = cstat + (c0 + c);  // element e at row[e * nchunks]
                fail = okey(lp_bound<NB>(b, row, row + d * nchunks, row[2 * d * nchunks],
                                         row[(2 * d + 1) * nchunks], nchunks)) < tau_hi;
            }
            const uint32_t m = (uint32_t)(__ballot(fail) >> (16 * g)) & 0xFFFFu;
            if (!done) {
                if (m) {
                    first = cb + (uint32_t)__builtin_ctz(m);
                    done = true;
                } else if (cb + 16 >= nc) {
                    done = true;
                }
            }
        }
        // objects to score: the chunks before the cut, the last one partial
        const uint32_t o = alive ? min(sz, first * PTA_CHUNK) : 0u;
        if (gl == 0 && p < L) {
            cut[p] = o;
            ch += (o + PTA_CHUNK - 1) / PTA_CHUNK;
            parts += o ? 1u : 0u;
            objs += o;
        }
    }
    for (int x = 32; x > 0; x >>= 1) {
        ch += __shfl_xor(ch, x, 64);
        parts += __shfl_xor(parts, x, 64);
        objs += __shfl_xor(objs, x, 64);
    }
    if (lane == 0) {
        wsum[wave][0] = ch;
        wsum[wave][1] = parts;
        wsum[wave][2] = objs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, bp = 0, bo = 0;
        for (int w = 0; w < PTA_BS / 64; w++) {
            a += wsum[w][0];
            bp += wsum[w][1];
            bo += wsum[w][2];
        }
        wchunks[blockIdx.x] = a;
        wstat[blockIdx.x] = make_uint2(bp, bo);
    }
}

// Scores the chunks of the partitions the cut pass kept.  Every workgroup
// first scans the G cut-workgroup chunk totals into LDS (E[g]: chunks
// before group g).  The waves split the chunks into equal contiguous
// ranges; a wave finds its group by a binary search of E, holds the
// group's cuts (lane j: partition g * R + j) and their chunk prefix in
// registers, and walks on, one chunk per step: 4 objects per lane, the
// queried columns loaded together.  Keys >= tau go to the workgroup's own
// slot region of the candidate buffer (an LDS counter; past the slot, the
// shared spill region), its count to cs.counts[workgroup].  Workgroup 0
// also reports the partitions and objects scored.
template <int M>
__global__ __launch_bounds__(PTA_BS) void pta_score_kernel(QCols q, uint32_t L, uint32_t R, uint32_t G,
                                                           const uint32_t *__restrict__ cut,
                                                           const uint32_t *__restrict__ wchunks,
                                                           const uint2 *__restrict__ wstat,
                                                           const uint32_t *__restrict__ poff,
                                                           const uint32_t *__restrict__ pid,
                                                           FteState *__restrict__ st, CandSet cs,
                                                           uint32_t *__restrict__ report) {
    __shared__ uint32_t E[PTA_CUT_GROUPS + 1];
    __shared__ uint32_t wtot[PTA_BS / 64];
    __shared__ uint32_t wrep[PTA_BS / 64][2];
    __shared__ uint32_t lcount;
    // the wave index made visibly uniform: the chunk cursor and its loads
    // (group offsets, cuts, partition offsets) are then scalar
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = __builtin_amdgcn_readfirstlane(tid / 64);
    // exclusive scan of the group totals: thread t owns groups [16t, 16t + 16)
    {
        uint32_t v[16], sum = 0, pp = 0, oo = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t gi = tid * 16 + j;
            v[j] = gi < G ? wchunks[gi] : 0u;
            sum += v[j];
        }
        if (blockIdx.x == 0) {
            // partitions and objects scored, for the report (fit 32 bits: positions do)
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t gi = tid * 16 + j;
                const uint2 w = gi < G ? wstat[gi] : make_uint2(0u, 0u);
                pp += w.x;
                oo += w.y;
            }
            for (int x = 32; x > 0; x >>= 1) {
                pp += __shfl_xor(pp, x, 64);
                oo += __shfl_xor(oo, x, 64);
            }
            if (lane == 0) {
                wrep[wave][0] = pp;
                wrep[wave][1] = oo;
            }
        }
        const uint32_t incl = wave_incl_scan(sum);
        if (lane == 63) wtot[wave] = incl;
        if (tid == 0) lcount = 0;
        __syncthreads();
        uint32_t run = incl - sum, all = 0;
        for (uint32_t w = 0; w < PTA_BS / 64; w++) {
            run += w < wave ? wtot[w] : 0u;
            all += wtot[w];
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t gi = tid * 16 + j;
            if (gi < G) E[gi] = run;
            run += v[j];
        }
        if (tid == 0) E[G] = all;  // the total
        if (blockIdx.x == 0 && tid == 0) {
            uint32_t a = 0, o = 0;
            for (int w = 0; w < PTA_BS / 64; w++) {
                a += wrep[w][0];
                o += wrep[w][1];
            }
            report[1] = a;
            report[2] = o;
            report[3] = 0;
        }
        __syncthreads();
    }
    const uint32_t total = E[G];
    const uint32_t nw = gridDim.x * (PTA_BS / 64), gw = blockIdx.x * (PTA_BS / 64) + wave;
    const uint64_t tau = st->tau;
    const uint32_t tau_hi = (uint32_t)(tau >> 32);
    const uint32_t per = (total + nw - 1) / nw;
    uint32_t c = gw * per;
    const uint32_t c_end = c + per < total ? c + per : total;
    uint64_t *my = cs.keys + (uint64_t)blockIdx.x * cs.slot;
    if (c < c_end) {
        // group of chunk c: the last g with E[g] <= c (E non-decreasing, E[0] = 0)
        uint32_t lo = 0, hi = G;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (E[mid] <= c) lo = mid;
            else hi = mid;
        }
        uint32_t grp = lo;
        uint32_t gbase = 0, lch = 0, lincl = 0;  // lane j: chunks of partition grp * R + j, inclusive prefix
        auto load_group = [&]() {
            gbase = E[grp];
            const uint32_t p = grp * R + lane;
            const uint32_t o = lane < R && p < L ? cut[p] : 0u;
            lch = (o + PTA_CHUNK - 1) / PTA_CHUNK;
            lincl = wave_incl_scan(lch);
        };
        load_group();
        for (; c < c_end; c++) {
            while (c >= E[grp + 1]) {
                grp++;
                if (c < E[grp + 1]) load_group();
            }
            // partition: the first lane whose inclusive prefix passes c
            const uint64_t pm = __ballot(gbase + lincl > c);
            const uint32_t j = (uint32_t)__builtin_ctzll(pm);
            const uint32_t pincl = (uint32_t)__shfl(lincl, (int)j, 64), pch = (uint32_t)__shfl(lch, (int)j, 64);
            const uint32_t p = grp * R + j, lc = c - (gbase + pincl - pch);
            const uint32_t objs_p = cut[p];  // objects to score (uniform load)
            const uint32_t first = lc * PTA_CHUNK + 4 * lane;
            const uint64_t pos = (uint64_t)poff[p] + first;
            f4 sc = {0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (M > 0) {
                f4 x[M];
#pragma unroll
                for (int jj = 0; jj < M; jj++) x[jj] = ld_stream(q.col[jj] + pos);
#pragma unroll
                for (int jj = 0; jj < M; jj++) {
                    f4 pr = x[jj] * q.w[jj];
                    sc = sc + pr;
                }
            } else {
                for (uint32_t jj = 0; jj < q.m; jj++) {
                    f4 pr = ld_stream(q.col[jj] + pos) * q.w[jj];
                    sc = sc + pr;
                }
            }
            const uint32_t nvalid = first < objs_p ? (objs_p - first < 4u ? objs_p - first : 4u) : 0u;
            uint64_t kk[4];
            uint32_t npass = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t ok = okey(sc[e]);
                uint64_t key = 0;
                if ((uint32_t)e < nvalid && ok >= tau_hi) {
                    key = rkey_from_okey(ok, pid[pos + e]);
                    if (key < tau) key = 0;
                }
                kk[e] = key;
                npass += key ? 1u : 0u;
            }
            if (__ballot(npass != 0)) {
                const uint32_t incl = wave_incl_scan(npass);
                const uint32_t wt = __shfl(incl, 63, 64);
                uint32_t cnt = 0;
                if (lane == 63) cnt = atomicAdd(&lcount, wt);
                cnt = __shfl(cnt, 63, 64);
                const uint32_t first_spill = cnt > cs.slot ? cnt : cs.slot;
                const uint32_t nspill = cnt + wt > first_spill ? cnt + wt - first_spill : 0u;
                uint32_t sbase = 0;
                if (nspill) {
                    if (lane == 63) sbase = atomicAdd(&st->spill, nspill);
                    sbase = __shfl(sbase, 63, 64);
                }
                uint32_t pos_k = cnt + incl - npass;
#pragma unroll
                for (int e = 0; e < 4; e++)
                    if (kk[e]) {
                        if (pos_k < cs.slot) {
                            my[pos_k] = kk[e];
                        } else {
                            const uint32_t sp = sbase + (pos_k - first_spill);
                            if (sp < cs.spill_cap) cs.spill[sp] = kk[e];
                        }
                        pos_k++;
                    }
            }
        }
    }
    __syncthreads();
    if (tid == 0) cs.counts[blockIdx.x] = lcount < cs.slot ? lcount : cs.slot;
}

// ---------------------------------------------------------------------------
// host: query
// ---------------------------------------------------------------------------
// bound -> seed -> chunk tasks -> chunk scoring -> final select, all on the
// context's stream; fo points at pinned host memory, so the caller only
// synchronises.
int pta_query(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, const FinalOut &fo) {
    PtaLayout &P = ctx->pta;
    if (!P.built) return TOPK_ERR_STATE;
    QCols q;
    int rc = build_qcols(ctx, P.pcols, P.ld, qq, weights, attr, q);
    if (rc) return rc;
    const BoundArgs b = make_bound_args(ctx->d, qq, weights, attr);

    Lane &Ln = ctx->lane[0];
    hipStream_t s = Ln.stream;
    const uint32_t L = P.nparts;
    const uint32_t nb = (L + PTA_BS - 1) / PTA_BS;
    hipLaunchKernelGGL(pta_bound_kernel, dim3(nb), dim3(PTA_BS), 0, s, b, L, P.psz, P.pmax, P.pmin, P.psum,
                       P.pminsum, P.pbound, P.bmax);
    // seed partitions are ranked by their own bounds when there are few of
    // them (so a small layout still seeds from several partitions), else by
    // the block maxima
    const bool own = L <= PTA_SEED_RANK;
    const uint64_t *rkeys = own ? P.pbound : P.bmax;
    const uint32_t nr = own ? L : nb;
    uint32_t T = (4 * k + PTA_SEED_OBJS - 1) / PTA_SEED_OBJS;
    if (T < PTA_MIN_SEED) T = PTA_MIN_SEED;
    if (T > PTA_MAX_SEED) T = PTA_MAX_SEED;
    if (T > nr) T = nr;
#define TK_SEED(MM)                                                                                              \
    hipLaunchKernelGGL(pta_seed_kernel<MM>, dim3(T), dim3(PTA_BS), 0, s, q, rkeys, nr, P.poff, P.psz, P.pid, k, \
                       Ln.blk, Ln.pre, Ln.st)
    TK_SWITCH_M(q.m, TK_SEED);
#undef TK_SEED
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    // cut pass: R partitions per workgroup (four per wave per round), at
    // most PTA_CUT_GROUPS workgroups
    uint32_t R = 16;
    while ((L + R - 1) / R > PTA_CUT_GROUPS) R += 16;
    if (R > 64) return TOPK_ERR_UNSUPPORTED;  // a scoring wave holds one group's cuts
    const uint32_t G = (L + R - 1) / R;
    if (ctx->d <= 8)
        hipLaunchKernelGGL(pta_cut_kernel<8>, dim3(G), dim3(PTA_BS), 0, s, b, L, R, P.pbound, P.poff, P.psz, P.cstat,
                           (uint64_t)P.nchunks, Ln.st, P.pcut, P.wchunks, P.wstat);
    else if (ctx->d <= 16)
        hipLaunchKernelGGL(pta_cut_kernel<16>, dim3(G), dim3(PTA_BS), 0, s, b, L, R, P.pbound, P.poff, P.psz,
                           P.cstat, (uint64_t)P.nchunks, Ln.st, P.pcut, P.wchunks, P.wstat);
    else
        hipLaunchKernelGGL(pta_cut_kernel<TOPK_MAX_DIMS>, dim3(G), dim3(PTA_BS), 0, s, b, L, R, P.pbound, P.poff,
                           P.psz, P.cstat, (uint64_t)P.nchunks, Ln.st, P.pcut, P.wchunks, P.wstat);
    // scoring: four resident 4-wave workgroups per CU, one slot region each
    const uint32_t sg = std::min((uint32_t)P.cus * 4, MAX_STREAM_GRID);
    CandSet cs;
    cs.keys = Ln.cand;
    cs.counts = Ln.counts;
    cs.nblk = sg;
    cs.slot = (uint32_t)(ctx->cand_cap / cs.nblk);
    cs.pre = nullptr;
    cs.spill = Ln.spill;
    cs.spill_cap = ctx->spill_cap;
#define TK_SCORE(MM)                                                                                          \
    hipLaunchKernelGGL(pta_score_kernel<MM>, dim3(sg), dim3(PTA_BS), 0, s, q, L, R, G, P.pcut, P.wchunks, P.wstat, \
                       P.poff, P.pid, Ln.st, cs, P.h_report)
    TK_SWITCH_M(q.m, TK_SCORE);
#undef TK_SCORE
    timing_end(ctx, s, "pta_stream", ev0);
    const uint32_t kreq = (uint32_t)(ctx->n < k ? ctx->n : k);
    FinalOut f = fo;
    f.cand = P.h_report;  // report[0]
    final_enqueue(ctx, Ln, cs, k, kreq, f, false);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

}  // namespace topk
