// Store preprocessing and the PTA entry points of the C ABI.
//
//  * normalisation: common/common.h:29-60 (per column, max starting at 0 and
//    min at FLT_MAX, x = (x - min) / (max - min) in binary32), called by
//    TPAc::init / PTA::init (cpu/TPAc.h:28-32, cpu/PTA.h:271).
//  * reorder_attr: gpu/reorder_attr.h:13-236 (per object, sort the d values
//    descending across the columns, in place); check_order invariant
//    gpu/GPA.h:96-114.
//  * PTA: layout build (pta.hip) and the pruned query.
#include <string.h>

#include <chrono>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace topk {

int pta_build(topk_ctx *ctx);
void pta_release(topk_ctx *ctx);
int pta_query(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, const FinalOut &fo);

// Per column max (starting at 0) and min (starting at FLT_MAX) as monotone
// keys, so atomicMax / atomicMin on uint32 implement float max / min, and
// the index of the column's first zero (either sign).  The monotone keys
// order -0 below +0, where the reference's sequential std::min keeps
// whichever zero it meets first; when the minimum is a zero, the host takes
// the sign of the first zero instead (nothing else depends on the order).
__global__ __launch_bounds__(256) void colminmax_kernel(const float *__restrict__ cols, uint64_t n, uint64_t ld,
                                                        uint32_t *__restrict__ kmax, uint32_t *__restrict__ kmin,
                                                        unsigned long long *__restrict__ zfirst) {
    const uint32_t m = blockIdx.y;
    const float *c = cols + (uint64_t)m * ld;
    float mx = 0.0f, mn = 3.40282347e+38f;
    uint64_t z = ~0ull;
    auto take = [&](float v, uint64_t i) {
        mx = mx < v ? v : mx;   // std::max(mmax, value)
        mn = v < mn ? v : mn;   // std::min(mmin, value)
        z = (v == 0.0f && i < z) ? i : z;
    };
    // 16-byte loads over whole quads (columns start 32 KiB aligned), then the
    // tail; the result is the same in any order (the cross-workgroup merge
    // takes the extremes of monotone keys and the smallest zero index)
    const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x, stride = (uint64_t)gridDim.x * 256;
    const uint64_t n4 = n / 4;
    for (uint64_t i = gid; i < n4; i += stride) {
        const float4 v = reinterpret_cast<const float4 *>(c)[i];
        take(v.x, 4 * i);
        take(v.y, 4 * i + 1);
        take(v.z, 4 * i + 2);
        take(v.w, 4 * i + 3);
    }
    for (uint64_t i = n4 * 4 + gid; i < n; i += stride) take(c[i], i);
    for (int o = 32; o > 0; o >>= 1) {
        const float a = __shfl_xor(mx, o, 64), b = __shfl_xor(mn, o, 64);
        mx = mx < a ? a : mx;
        mn = b < mn ? b : mn;
    }
    z = wave_min_u64(z);
    // one global atomic per statistic and workgroup, on the monotone keys
    __shared__ uint32_t wmx[4], wmn[4];
    __shared__ uint64_t wz[4];
    if (lane_id() == 0) {
        wmx[threadIdx.x / 64] = okey(mx);
        wmn[threadIdx.x / 64] = okey(mn);
        wz[threadIdx.x / 64] = z;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = wmx[0], b = wmn[0];
        uint64_t zz = wz[0];
        for (int w = 1; w < 4; w++) {
            a = wmx[w] > a ? wmx[w] : a;
            b = wmn[w] < b ? wmn[w] : b;
            zz = wz[w] < zz ? wz[w] : zz;
        }
        atomicMax(&kmax[m], a);
        atomicMin(&kmin[m], b);
        if (zz != ~0ull) atomicMin(&zfirst[m], (unsigned long long)zz);
    }
}

// Per column minimum and range (max - min), and whether every column can
// be normalised (computed on the device: no host round trip between the
// statistics and the writes).
struct NormParams {
    float mn[TOPK_MAX_DIMS];
    float mm[TOPK_MAX_DIMS];
    uint32_t ok;
};

// One wave; lane m < d takes column m.  A minimum of zero is read back from
// the column's first zero, so its sign is the stored one.
__global__ void norm_params_kernel(const float *__restrict__ cols, uint64_t ld, const uint32_t *__restrict__ k,
                                   const unsigned long long *__restrict__ z, uint32_t d, NormParams *__restrict__ p) {
    const uint32_t m = lane_id();
    bool good = true;
    if (m < d) {
        float mn = okey_inv(k[d + m]);
        if (mn == 0.0f && z[m] != ~0ull) mn = cols[(uint64_t)m * ld + z[m]];
        const float mm = okey_inv(k[m]) - mn;
        p->mn[m] = mn;
        p->mm[m] = mm;
        good = mm > 0.0f && mm <= 3.40282347e+38f;
    }
    const uint64_t bad = __ballot(!good);
    if (m == 0) p->ok = bad == 0ull ? 1u : 0u;
}

// Writes nothing unless every column can be normalised (p->ok).
__global__ void normalize_kernel(float *__restrict__ cols, uint64_t n, uint64_t ld, const NormParams *__restrict__ p) {
    if (!p->ok) return;
    const uint32_t m = blockIdx.y;
    const float mn = p->mn[m], mm = p->mm[m];
    float *c = cols + (uint64_t)m * ld;
    const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x, stride = (uint64_t)gridDim.x * 256;
    const uint64_t n4 = n / 4;
    for (uint64_t i = gid; i < n4; i += stride) {
        float4 v = reinterpret_cast<float4 *>(c)[i];
        v.x = (v.x - mn) / mm;
        v.y = (v.y - mn) / mm;
        v.z = (v.z - mn) / mm;
        v.w = (v.w - mn) / mm;
        reinterpret_cast<float4 *>(c)[i] = v;
    }
    for (uint64_t i = n4 * 4 + gid; i < n; i += stride) c[i] = (c[i] - mn) / mm;
}

// Per object, its d values sorted descending across the columns
// (gpu/reorder_attr.h:13-236).  A stable odd-even transposition network over
// D register slots (D >= d; slots past d hold -inf and stay last): adjacent
// pairs swap only when strictly out of order, so equal values (0 and -0)
// keep their column order, as in an insertion sort.
template <int D>
__global__ __launch_bounds__(256) void reorder_attr_kernel(float *__restrict__ cols, uint64_t n, uint32_t d,
                                                           uint64_t ld) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        float v[D];
#pragma unroll
        for (int m = 0; m < D; m++) v[m] = (uint32_t)m < d ? cols[(uint64_t)m * ld + i] : -INFINITY;
#pragma unroll
        for (int r = 0; r < D; r++) {
#pragma unroll
            for (int a = r & 1; a + 1 < D; a += 2) {
                const float x = v[a], y = v[a + 1];
                const bool sw = x < y;
                v[a] = sw ? y : x;
                v[a + 1] = sw ? x : y;
            }
        }
#pragma unroll
        for (int m = 0; m < D; m++)
            if ((uint32_t)m < d) cols[(uint64_t)m * ld + i] = v[m];
    }
}

// Column maxima start at 0 (common/common.h:36), minima at FLT_MAX, first
// zero indices at "none".
__global__ void minmax_init_kernel(uint32_t *k, unsigned long long *z, uint32_t d) {
    const uint32_t m = threadIdx.x;
    if (m < d) {
        k[m] = okey(0.0f);
        k[d + m] = okey(3.40282347e+38f);
        z[m] = ~0ull;
    }
}

// The store holds only finite values (checked when it is built), and
// mx >= every value >= mn, so 0 <= x - mn <= mx - mn = mm for every x of a
// column (binary32 subtraction is monotone): every (x - mn) / mm is finite
// exactly when mm is finite and > 0.  The ranges are checked on the device
// before anything is written, so a column that cannot be normalised
// (constant, or a range past FLT_MAX) leaves the whole store untouched.
static int normalize_store(topk_ctx *ctx) {
    hipStream_t s = ctx->stream;
    const uint32_t d = ctx->d;
    unsigned long long *z = nullptr;
    TK_HIP(hipMalloc(&z, sizeof(unsigned long long) * TOPK_MAX_DIMS + sizeof(uint32_t) * 2 * TOPK_MAX_DIMS +
                             sizeof(NormParams)));
    uint32_t *k = reinterpret_cast<uint32_t *>(z + TOPK_MAX_DIMS);
    NormParams *np = reinterpret_cast<NormParams *>(k + 2 * TOPK_MAX_DIMS);
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, s, ev0);
    hipLaunchKernelGGL(minmax_init_kernel, dim3(1), dim3(TOPK_MAX_DIMS), 0, s, k, z, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(colminmax_kernel, dim3(512, d), dim3(256), 0, s, ctx->cols, ctx->n, ctx->ld, k, k + d, z);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(norm_params_kernel, dim3(1), dim3(64), 0, s, ctx->cols, ctx->ld, k, z, d, np);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(normalize_kernel, dim3(2048, d), dim3(256), 0, s, ctx->cols, ctx->n, ctx->ld, np);
        e = hipGetLastError();
    }
    timing_end(ctx, s, "normalize", ev0);
    uint32_t ok = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&ok, &np->ok, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(z);
    TK_HIP(e);
    return ok ? TOPK_OK : TOPK_ERR_NONFINITE;
}

}  // namespace topk

using namespace topk;

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

int topk_init(topk_ctx *ctx, uint32_t flags) {
    if (!ctx) return TOPK_ERR_INVALID;
    if (flags & ~(TOPK_INIT_NORMALIZE | TOPK_INIT_PTA)) return TOPK_ERR_INVALID;
    const double t0 = now_ms();
    int rc = TOPK_OK;
    if (flags & TOPK_INIT_NORMALIZE) {
        pta_release(ctx);  // a layout built from the old values would be stale
        ctx->store_gen++;
        rc = normalize_store(ctx);
        if (rc) return rc;
    }
    if (flags & TOPK_INIT_PTA) {
        rc = pta_build(ctx);
        if (rc) return rc;
    }
    ctx->stats.tt_init = now_ms() - t0;
    return TOPK_OK;
}

int topk_set_partition_size(topk_ctx *ctx, uint32_t objects) {
    if (!ctx || objects < 1024 || objects > 65536 || objects % 1024) return TOPK_ERR_INVALID;
    ctx->pta.psize = objects;
    return TOPK_OK;
}

int topk_reorder_attr(topk_ctx *ctx) {
    if (!ctx) return TOPK_ERR_INVALID;
    pta_release(ctx);
    ctx->store_gen++;
    hipEvent_t ev0 = nullptr;
    timing_begin(ctx, ctx->stream, ev0);
    // one object per work-item, grid-stride past 2^28 workgroups
    const uint64_t nb = (ctx->n + 255) / 256;
    const dim3 grid((uint32_t)(nb < (1u << 20) ? nb : (1u << 20)));
    if (ctx->d <= 8)
        hipLaunchKernelGGL(reorder_attr_kernel<8>, grid, dim3(256), 0, ctx->stream, ctx->cols, ctx->n, ctx->d, ctx->ld);
    else if (ctx->d <= 16)
        hipLaunchKernelGGL(reorder_attr_kernel<16>, grid, dim3(256), 0, ctx->stream, ctx->cols, ctx->n, ctx->d,
                           ctx->ld);
    else
        hipLaunchKernelGGL(reorder_attr_kernel<TOPK_MAX_DIMS>, grid, dim3(256), 0, ctx->stream, ctx->cols, ctx->n,
                           ctx->d, ctx->ld);
    timing_end(ctx, ctx->stream, "reorder_attr", ev0);
    TK_HIP(hipGetLastError());
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TOPK_OK;
}

int topk_pta(topk_ctx *ctx, uint32_t k, uint8_t qq, const float *weights, const uint8_t *attr, uint64_t *out_ids,
             float *out_scores, float *out_threshold) {
    if (!ctx || k < 1) return TOPK_ERR_INVALID;
    if (k > TOPK_MAX_K) return TOPK_ERR_UNSUPPORTED;
    if (!ctx->pta.built) return TOPK_ERR_STATE;
    const double t0 = now_ms();
    // the final select writes the pinned result block and report directly
    uint64_t *h_ids = reinterpret_cast<uint64_t *>(ctx->h_res + ctx->res.ids_off);
    float *h_sc = reinterpret_cast<float *>(ctx->h_res + ctx->res.scores_off);
    float *h_thr = reinterpret_cast<float *>(ctx->h_res + ctx->res.thr_off);
    uint32_t *h_status = reinterpret_cast<uint32_t *>(ctx->h_res + ctx->res.status_off);
    int rc = pta_query(ctx, k, qq, weights, attr, FinalOut{h_ids, h_sc, h_thr, h_status});
    if (rc) return rc;
    Lane &L = ctx->lane[0];
    TK_HIP(hipStreamSynchronize(L.stream));
    const uint32_t *rep = ctx->pta.h_report;
    const uint64_t cand = rep[0], scored = rep[1], objs = (uint64_t)rep[2] | ((uint64_t)rep[3] << 32);
    const uint32_t status = *reinterpret_cast<const uint32_t *>(ctx->h_res + ctx->res.status_off);
    if (status) {
        // candidate storage overflowed (pathological data): the exact FTE path answers
        TK_HIP(hipMemsetAsync(&L.st->status, 0, sizeof(uint32_t), L.stream));
        rc = topk_fte(ctx, k, qq, weights, attr, out_ids, out_scores, out_threshold);
        ctx->stats.fallback = 1;
        ctx->stats.partitions_scored = ctx->pta.nparts;
        ctx->stats.partitions_total = ctx->pta.nparts;
        return rc;
    }
    const uint64_t *ids = reinterpret_cast<const uint64_t *>(ctx->h_res + ctx->res.ids_off);
    const float *sc = reinterpret_cast<const float *>(ctx->h_res + ctx->res.scores_off);
    if (out_ids) memcpy(out_ids, ids, sizeof(uint64_t) * k);
    if (out_scores) memcpy(out_scores, sc, sizeof(float) * k);
    if (out_threshold) *out_threshold = *reinterpret_cast<const float *>(ctx->h_res + ctx->res.thr_off);
    ctx->stats.candidates = cand;
    ctx->stats.partitions_scored = scored;
    ctx->stats.partitions_total = ctx->pta.nparts;
    ctx->stats.fallback = 0;
    ctx->stats.retried = 0;
    ctx->stats.fte_fallback = 0;
    ctx->stats.queries = 1;
    ctx->stats.ms_last = now_ms() - t0;
    ctx->stats.tuple_count = objs;
    ctx->stats.terms = objs * qq;
    return TOPK_OK;
}

int topk_pta_layout(topk_ctx *ctx, uint32_t *nparts, uint64_t *npos) {
    if (!ctx) return TOPK_ERR_INVALID;
    if (!ctx->pta.built) return TOPK_ERR_STATE;
    if (nparts) *nparts = ctx->pta.nparts;
    if (npos) *npos = ctx->pta.npos;
    return TOPK_OK;
}

int topk_pta_read_layout(topk_ctx *ctx, uint32_t *pid, uint32_t *poff, uint32_t *psz, float *pmax, float *pmin) {
    if (!ctx) return TOPK_ERR_INVALID;
    const PtaLayout &P = ctx->pta;
    if (!P.built) return TOPK_ERR_STATE;
    hipStream_t s = ctx->stream;
    if (pid) TK_HIP(hipMemcpyAsync(pid, P.pid, sizeof(uint32_t) * P.npos, hipMemcpyDeviceToHost, s));
    if (poff) TK_HIP(hipMemcpyAsync(poff, P.poff, sizeof(uint32_t) * (P.nparts + 1), hipMemcpyDeviceToHost, s));
    if (psz) TK_HIP(hipMemcpyAsync(psz, P.psz, sizeof(uint32_t) * P.nparts, hipMemcpyDeviceToHost, s));
    if (pmax) TK_HIP(hipMemcpyAsync(pmax, P.pmax, sizeof(float) * P.nparts * ctx->d, hipMemcpyDeviceToHost, s));
    if (pmin) TK_HIP(hipMemcpyAsync(pmin, P.pmin, sizeof(float) * P.nparts * ctx->d, hipMemcpyDeviceToHost, s));
    TK_HIP(hipStreamSynchronize(s));
    return TOPK_OK;
}

}  // extern "C"
