// Learned k-th score predictor for the batch path.
//
// The reference's PTA prunes with a conservative threshold: the running k-th
// score, which starts at -inf and only rises as partitions are scanned
// (cpu/PTA.h:357-422).  Here a small MLP predicts a query's k-th score up
// front from (query, data sketch); the batch path lowers the prediction by a
// calibrated margin and uses it as the query's threshold tau in its fused
// score + prune + compact pass.  Exactness never depends on the network:
// fewer than k survivors re-run the query exactly with a lower threshold in
// the batch's second pass, and a query still open after it goes to the plain
// exact path (batch.hip).
//
//  * network: 96 -> 64 -> 64 -> 1, ReLU, FP32.  Parameter layout W1[64][96],
//    b1[64], W2[64][64], b2[64], W3[64], b3 (10433 floats).
//  * features (mlp_features): 16 weights by attribute id normalised by
//    sum |W|, log2 k, per-attribute quantiles {0.5, 0.9, 0.99, 0.999} of a 1%
//    sample (the data sketch), and 15 derived terms (weighted quantile sums,
//    sizes).  Target: k-th score / sum |W|.
//  * training: labels are the k-th scores the exact batch path (seed mode)
//    returns for the training queries; mini-batch Adam on the MSE with
//    hand-written forward / backward / update kernels (mlp_pass_kernel,
//    mlp_adam_kernel); gradients match the FP64 oracle restatement.
//  * calibration: margin = the (1 - target_fallback) quantile of
//    (prediction - label) on held-out queries.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace topk {

constexpr int MLP_IN = 96, MLP_H = 64;
constexpr uint32_t OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H,
                   OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H,
                   MLP_NP = OFF_B3 + 1;
static_assert(MLP_NP == 10433, "parameter count");
constexpr int LRN_MAX_D = 16;
constexpr int MLP_ROWS = 64;   // samples per forward / backward workgroup
constexpr int MLP_BS = 256;
constexpr int TAU_BS = 256;    // inference workgroup: 4 waves, one query per wave at a time

struct Sketch {
    float q[LRN_MAX_D][4];     // quantiles 0.5, 0.9, 0.99, 0.999 of a 1% sample, per attribute
    float cmax[LRN_MAX_D];     // sample maximum per attribute
    float log2n;
};

struct LearnedState {
    bool ready = false;
    uint64_t version = 0;      // bumped by every install
    float *d_params = nullptr; // installed predictor
    float margin = 0.0f;       // in units of score / sum |W|
    Sketch sk{};
    bool sk_valid = false;
    // The store saturates: training labels at the largest score the data
    // allows (many objects tied there, e.g. values clipped at a bound).  A
    // score threshold admits every tie, so queries predicted near their
    // maximum take an exact seed threshold (a key) instead.
    bool saturated = false;
};

static LearnedState *lstate(topk_ctx *ctx) {
    if (!ctx->learned) ctx->learned = new LearnedState();
    return static_cast<LearnedState *>(ctx->learned);
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// f[0..96) for query Q; *scale = sum |W| (W aggregated per attribute).
__device__ void mlp_features(const BatchQuery &Q, const Sketch &S, float *f, float *scale) {
    float wa[LRN_MAX_D];
#pragma unroll
    for (int a = 0; a < LRN_MAX_D; a++) wa[a] = 0.0f;
    for (uint32_t j = 0; j < Q.m; j++) {
        const uint32_t at = Q.attr[j];
#pragma unroll
        for (int a = 0; a < LRN_MAX_D; a++) wa[a] = (uint32_t)a == at ? wa[a] + Q.w[j] : wa[a];
    }
    float sabs = 0.0f;
#pragma unroll
    for (int a = 0; a < LRN_MAX_D; a++) sabs += fabsf(wa[a]);
    const float inv = sabs > 0.0f ? 1.0f / sabs : 0.0f;
    float s1 = 0.0f, s2 = 0.0f, mx = -1.0f, sc = 0.0f, sq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < LRN_MAX_D; a++) {
        const float wn = wa[a] * inv;
        f[a] = wn;
        s1 += wn;
        s2 += wn * wn;
        mx = wn > mx ? wn : mx;
        sc += wn * S.cmax[a];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f[17 + 4 * a + t] = S.q[a][t];
            sq[t] += wn * S.q[a][t];
        }
    }
    const float lk = log2f((float)Q.k);
    f[16] = lk * 0.1f;
    f[81] = (float)Q.m * 0.0625f;
    f[82] = s1;
    f[83] = s2;
    f[84] = mx;
    f[85] = S.log2n * 0.03125f;
    f[86] = (lk - S.log2n) * 0.03125f;
#pragma unroll
    for (int t = 0; t < 4; t++) f[87 + t] = sq[t];
    f[91] = sc;
#pragma unroll
    for (int i = 92; i < MLP_IN; i++) f[i] = 0.0f;
    *scale = sabs;
}

__global__ void mlp_feat_kernel(const BatchQuery *__restrict__ qs, uint32_t nq, Sketch S, float *__restrict__ X,
                                float *__restrict__ scale) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    float f[MLP_IN], s;
    mlp_features(qs[q], S, f, &s);
    for (int i = 0; i < MLP_IN; i++) X[(uint64_t)q * MLP_IN + i] = f[i];
    scale[q] = s;
}

// ---------------------------------------------------------------------------
// forward / backward over a block of up to 64 samples (LDS resident)
// ---------------------------------------------------------------------------
struct MlpLds {
    float W1T[MLP_IN][MLP_H];
    float W2T[MLP_H][MLP_H];
    float W2[MLP_H][MLP_H];
    float b1[MLP_H], b2[MLP_H], W3[MLP_H];
    float b3;
    float X[MLP_ROWS][MLP_IN];
    float H1[MLP_ROWS][MLP_H];
    float H2[MLP_ROWS][MLP_H];
    float G1[MLP_ROWS][MLP_H];
    float G2[MLP_ROWS][MLP_H];
    float out[MLP_ROWS], go[MLP_ROWS], e2[MLP_ROWS];
};

// Forward of rows rows[r0 + b] (b < nr; rows == nullptr: r0 + b) and, when
// G is non-null, the gradient of sum_b (out_b - y_b)^2 / total into
// G[blockIdx.x * MLP_NP ...] and the squared errors' sum into lsum[blockIdx.x].
template <bool BACKWARD>
__global__ __launch_bounds__(MLP_BS) void mlp_pass_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                          const float *__restrict__ Y,
                                                          const uint32_t *__restrict__ rows, uint32_t r0,
                                                          uint32_t nrows, float inv_total, float *__restrict__ out,
                                                          float *__restrict__ G, float *__restrict__ lsum) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MlpLds &L = *reinterpret_cast<MlpLds *>(lds_raw);
    const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid / 64;
    const uint32_t base = r0 + blockIdx.x * MLP_ROWS;
    const uint32_t nr = base < nrows ? (nrows - base < (uint32_t)MLP_ROWS ? nrows - base : (uint32_t)MLP_ROWS) : 0u;
    for (uint32_t o = tid; o < MLP_H * MLP_IN; o += MLP_BS) {
        const uint32_t j = o / MLP_IN, i = o % MLP_IN;
        L.W1T[i][j] = P[OFF_W1 + o];
    }
    for (uint32_t o = tid; o < MLP_H * MLP_H; o += MLP_BS) {
        const float w = P[OFF_W2 + o];
        L.W2[o / MLP_H][o % MLP_H] = w;
        L.W2T[o % MLP_H][o / MLP_H] = w;
    }
    if (tid < MLP_H) {
        L.b1[tid] = P[OFF_B1 + tid];
        L.b2[tid] = P[OFF_B2 + tid];
        L.W3[tid] = P[OFF_W3 + tid];
    }
    if (tid == 0) L.b3 = P[OFF_B3];
    for (uint32_t o = tid; o < MLP_ROWS * MLP_IN; o += MLP_BS) {
        const uint32_t b = o / MLP_IN, i = o % MLP_IN;
        float v = 0.0f;
        if (b < nr) {
            const uint32_t r = rows ? rows[base + b] : base + b;
            v = X[(uint64_t)r * MLP_IN + i];
        }
        L.X[b][i] = v;
    }
    __syncthreads();
    for (uint32_t o = tid; o < MLP_ROWS * MLP_H; o += MLP_BS) {
        const uint32_t b = o / MLP_H, j = o % MLP_H;
        float s = L.b1[j];
        for (int i = 0; i < MLP_IN; i++) s += L.X[b][i] * L.W1T[i][j];
        L.H1[b][j] = s > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    for (uint32_t o = tid; o < MLP_ROWS * MLP_H; o += MLP_BS) {
        const uint32_t b = o / MLP_H, j = o % MLP_H;
        float s = L.b2[j];
        for (int i = 0; i < MLP_H; i++) s += L.H1[b][i] * L.W2T[i][j];
        L.H2[b][j] = s > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    for (uint32_t b = wv; b < MLP_ROWS; b += MLP_BS / 64) {
        float s = L.W3[lane] * L.H2[b][lane];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) {
            const float y = s + L.b3;
            L.out[b] = y;
            if (b < nr) {
                const uint32_t r = rows ? rows[base + b] : base + b;
                if (out) out[base + b] = y;
                const float e = BACKWARD ? y - Y[r] : 0.0f;
                L.go[b] = 2.0f * e * inv_total;
                L.e2[b] = e * e;
            } else {
                L.go[b] = 0.0f;
                L.e2[b] = 0.0f;
            }
        }
    }
    if constexpr (!BACKWARD) return;
    __syncthreads();
    for (uint32_t o = tid; o < MLP_ROWS * MLP_H; o += MLP_BS) {
        const uint32_t b = o / MLP_H, j = o % MLP_H;
        L.G2[b][j] = L.H2[b][j] > 0.0f ? L.go[b] * L.W3[j] : 0.0f;
    }
    __syncthreads();
    for (uint32_t o = tid; o < MLP_ROWS * MLP_H; o += MLP_BS) {
        const uint32_t b = o / MLP_H, i = o % MLP_H;
        float s = 0.0f;
        for (int j = 0; j < MLP_H; j++) s += L.G2[b][j] * L.W2[j][i];
        L.G1[b][i] = L.H1[b][i] > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    float *g = G + (uint64_t)blockIdx.x * MLP_NP;
    for (uint32_t o = tid; o < MLP_H * MLP_IN; o += MLP_BS) {
        const uint32_t j = o / MLP_IN, i = o % MLP_IN;
        float s = 0.0f;
        for (int b = 0; b < MLP_ROWS; b++) s += L.G1[b][j] * L.X[b][i];
        g[OFF_W1 + o] = s;
    }
    for (uint32_t o = tid; o < MLP_H * MLP_H; o += MLP_BS) {
        const uint32_t j = o / MLP_H, i = o % MLP_H;
        float s = 0.0f;
        for (int b = 0; b < MLP_ROWS; b++) s += L.G2[b][j] * L.H1[b][i];
        g[OFF_W2 + o] = s;
    }
    if (tid < MLP_H) {
        float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (int b = 0; b < MLP_ROWS; b++) {
            s1 += L.G1[b][tid];
            s2 += L.G2[b][tid];
            s3 += L.go[b] * L.H2[b][tid];
        }
        g[OFF_B1 + tid] = s1;
        g[OFF_B2 + tid] = s2;
        g[OFF_W3 + tid] = s3;
    } else if (tid == MLP_H) {
        float s = 0.0f, l = 0.0f;
        for (int b = 0; b < MLP_ROWS; b++) {
            s += L.go[b];
            l += L.e2[b];
        }
        g[OFF_B3] = s;
        lsum[blockIdx.x] = l;
    }
}

// Sums the nb partial gradients; Adam step when lr > 0, else writes the sum.
__global__ void mlp_adam_kernel(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                const float *__restrict__ G, uint32_t nb, float lr, float b1, float b2, float eps,
                                float bc1, float bc2, float *__restrict__ gsum) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nb; b++) g += G[(uint64_t)b * MLP_NP + p];
    if (gsum) gsum[p] = g;
    if (lr <= 0.0f) return;
    const float m = b1 * M[p] + (1.0f - b1) * g;
    const float v = b2 * V[p] + (1.0f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    const float mh = m / bc1, vh = v / bc2;
    P[p] = P[p] - lr * mh / (sqrtf(vh) + eps);
}

// Threshold per batch query: tau = (prediction - margin) * sum |W| (or
// -inf when the weights are all zero).  pred (optional) gets the unlowered
// prediction in score units.  One wave per query: every lane builds the
// feature vector in registers, lane j owns hidden unit j of both layers and
// the second layer reads the first through lane shuffles (no barrier after
// the parameter load, so waves may take different numbers of queries).
__global__ __launch_bounds__(TAU_BS) void mlp_tau_kernel(const float *__restrict__ P, BatchQuery *__restrict__ qs,
                                                         uint32_t nq, Sketch S, float margin,
                                                         float *__restrict__ pred, float retry_width,
                                                         int saturated) {
    // Lane j owns hidden unit j of both layers: its rows of W1 and W2 (stored
    // row-major, row j = unit j) come straight from memory as 16-byte loads,
    // all in flight at once -- no LDS staging, no barrier -- and the second
    // layer reads the first through lane shuffles.
    const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid / 64;
    const uint32_t nw = gridDim.x * (TAU_BS / 64);
    const uint32_t q0 = blockIdx.x * (TAU_BS / 64) + wv;
    if (q0 >= nq) return;
    float w1[MLP_IN], w2[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_IN / 4; i++) {
        const float4 v = reinterpret_cast<const float4 *>(P + OFF_W1 + lane * MLP_IN)[i];
        w1[4 * i] = v.x;
        w1[4 * i + 1] = v.y;
        w1[4 * i + 2] = v.z;
        w1[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < MLP_H / 4; i++) {
        const float4 v = reinterpret_cast<const float4 *>(P + OFF_W2 + lane * MLP_H)[i];
        w2[4 * i] = v.x;
        w2[4 * i + 1] = v.y;
        w2[4 * i + 2] = v.z;
        w2[4 * i + 3] = v.w;
    }
    const float b1 = P[OFF_B1 + lane], b2 = P[OFF_B2 + lane], w3 = P[OFF_W3 + lane], b3 = P[OFF_B3];
    for (uint32_t q = q0; q < nq; q += nw) {
        float f[MLP_IN], sc;
        mlp_features(qs[q], S, f, &sc);
        float s = b1;
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) s += f[i] * w1[i];
        const float h1 = s > 0.0f ? s : 0.0f;
        float t = b2;
#pragma unroll
        for (int i = 0; i < MLP_H; i++) t += __shfl(h1, i, 64) * w2[i];
        t = t > 0.0f ? t : 0.0f;
        float o = w3 * t;
        for (int x = 32; x > 0; x >>= 1) o += __shfl_xor(o, x, 64);
        if (lane == 0) {
            const float y = o + b3;
            qs[q].tau = sc > 0.0f ? (y - margin) * sc : -INFINITY;
            // a miss reruns retry_width margins below the prediction, or
            // (retry_width < 0) with a seed threshold
            const float wide = retry_width * fmaxf(fabsf(margin), 1e-3f);  // strictly below tau
            qs[q].tau_retry = retry_width < 0.0f ? INFINITY : (sc > 0.0f ? (y - wide) * sc : -INFINITY);
            // within two margins of the query's largest score (feature 91):
            // ties likely, seed this query
            qs[q].seed = saturated && sc > 0.0f && y + 2.0f * fabsf(margin) >= f[91] ? 1u : 0u;
            if (pred) pred[q] = y * sc;
        }
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr size_t MLP_LDS = sizeof(MlpLds);

__global__ void sketch_sample_kernel(const float *__restrict__ cols, uint64_t ld, uint32_t d, uint64_t n, uint64_t s,
                                     float *__restrict__ out) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= s) return;
    const uint64_t i = j * n / s;
    for (uint32_t m = 0; m < d; m++) out[(uint64_t)m * s + j] = cols[(uint64_t)m * ld + i];
}

// Per-attribute quantiles of a 1% sample of the store (at least 4096 objects).
static int build_sketch(topk_ctx *ctx, Sketch &S) {
    const uint64_t n = ctx->n;
    uint64_t s = n / 100;
    if (s < 4096) s = n < 4096 ? n : 4096;
    std::vector<float> h((size_t)s * ctx->d);
    float *dv = nullptr;
    TK_HIP(hipMalloc(&dv, sizeof(float) * h.size()));
    hipLaunchKernelGGL(sketch_sample_kernel, dim3((uint32_t)((s + 255) / 256)), dim3(256), 0, ctx->stream, ctx->cols,
                       ctx->ld, ctx->d, n, s, dv);
    hipError_t e = hipMemcpyAsync(h.data(), dv, sizeof(float) * h.size(), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dv);
    TK_HIP(e);
    memset(&S, 0, sizeof(S));
    static const double qs[4] = {0.5, 0.9, 0.99, 0.999};
    for (uint32_t m = 0; m < ctx->d; m++) {
        float *c = h.data() + (size_t)m * s;
        for (int t = 0; t < 4; t++) {
            const size_t r = (size_t)(qs[t] * (double)(s - 1));
            std::nth_element(c, c + r, c + s);
            S.q[m][t] = c[r];
        }
        S.cmax[m] = *std::max_element(c, c + s);
    }
    S.log2n = (float)std::log2((double)n);
    return TOPK_OK;
}

// Device buffers of one training / evaluation run.
struct MlpBuffers {
    float *X = nullptr, *Xs = nullptr, *Y = nullptr, *scale = nullptr, *P = nullptr, *M = nullptr, *V = nullptr,
          *G = nullptr, *lsum = nullptr, *out = nullptr;
    uint32_t *rows = nullptr;
    BatchQuery *q = nullptr;
    ~MlpBuffers() {
        hipFree(X);
        hipFree(Xs);
        hipFree(Y);
        hipFree(scale);
        hipFree(P);
        hipFree(M);
        hipFree(V);
        hipFree(G);
        hipFree(lsum);
        hipFree(out);
        hipFree(rows);
        hipFree(q);
    }
};

static int set_pass_lds() {
    static int done = 0;
    if (done) return TOPK_OK;
    TK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(mlp_pass_kernel<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)MLP_LDS));
    TK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(mlp_pass_kernel<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)MLP_LDS));
    done = 1;
    return TOPK_OK;
}

// Forward (+ backward) of nrows rows [r0, r0 + nrows) of X (through rows[]
// when given).  Gradient sum -> gsum (optional), squared-error sum -> *l2.
static int mlp_pass(hipStream_t s, const float *P, const float *X, const float *Y, const uint32_t *rows, uint32_t r0,
                    uint32_t nrows, float inv_total, float *out, float *G, float *lsum, bool backward) {
    const uint32_t nb = (nrows + MLP_ROWS - 1) / MLP_ROWS;
    if (!nb) return TOPK_OK;
    if (backward)
        hipLaunchKernelGGL(mlp_pass_kernel<true>, dim3(nb), dim3(MLP_BS), MLP_LDS, s, P, X, Y, rows, r0, r0 + nrows,
                           inv_total, out, G, lsum);
    else
        hipLaunchKernelGGL(mlp_pass_kernel<false>, dim3(nb), dim3(MLP_BS), MLP_LDS, s, P, X, Y, rows, r0,
                           r0 + nrows, inv_total, out, G, lsum);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

bool learned_ready(const topk_ctx *ctx) {
    const LearnedState *L = static_cast<const LearnedState *>(ctx->learned);
    return L && L->ready;
}

uint64_t learned_version(const topk_ctx *ctx) {
    const LearnedState *L = static_cast<const LearnedState *>(ctx->learned);
    return L ? L->version : 0;
}

bool learned_saturated(const topk_ctx *ctx) {
    const LearnedState *L = static_cast<const LearnedState *>(ctx->learned);
    return L && L->ready && L->saturated;
}

int learned_tau_enqueue(topk_ctx *ctx, hipStream_t s, BatchQuery *q, uint32_t nq) {
    LearnedState *L = static_cast<LearnedState *>(ctx->learned);
    if (!L || !L->ready) return TOPK_ERR_STATE;
    uint32_t g = (nq + TAU_BS / 64 - 1) / (TAU_BS / 64);
    if (g > 1024) g = 1024;
    const float width =
        ctx->tune.batch_retry == 2 ? -1.0f : (ctx->tune.batch_retry_width ? (float)ctx->tune.batch_retry_width : 3.0f);
    hipLaunchKernelGGL(mlp_tau_kernel, dim3(g), dim3(TAU_BS), 0, s, L->d_params, q, nq, L->sk, L->margin,
                       (float *)nullptr, width, L->saturated ? 1 : 0);
    TK_HIP(hipGetLastError());
    return TOPK_OK;
}

void learned_free(topk_ctx *ctx) {
    LearnedState *L = static_cast<LearnedState *>(ctx->learned);
    if (!L) return;
    hipFree(L->d_params);
    delete L;
    ctx->learned = nullptr;
}

static int install(topk_ctx *ctx, const float *d_src, float margin) {
    batch_quiesce(ctx);  // batches in flight read the parameters
    LearnedState *L = lstate(ctx);
    if (!L->d_params) TK_HIP(hipMalloc(&L->d_params, sizeof(float) * MLP_NP));
    TK_HIP(hipMemcpyAsync(L->d_params, d_src, sizeof(float) * MLP_NP, hipMemcpyDefault, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    L->margin = margin;
    L->ready = true;
    L->version++;
    return TOPK_OK;
}

// Stages queries as the batch path does and builds their features on the
// device: X[nq][96], scale[nq].
static int features(topk_ctx *ctx, const Sketch &S, uint32_t nq, const topk_query *qs, MlpBuffers &B) {
    std::vector<BatchQuery> hq(nq);
    for (uint32_t i = 0; i < nq; i++) {
        const int rc = make_batch_query(ctx, qs[i], hq[i]);
        if (rc) return rc;
    }
    TK_HIP(hipMalloc(&B.q, sizeof(BatchQuery) * nq));
    TK_HIP(hipMalloc(&B.X, sizeof(float) * MLP_IN * nq));
    TK_HIP(hipMalloc(&B.scale, sizeof(float) * nq));
    TK_HIP(hipMemcpyAsync(B.q, hq.data(), sizeof(BatchQuery) * nq, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mlp_feat_kernel, dim3((nq + 255) / 256), dim3(256), 0, ctx->stream, B.q, nq, S, B.X, B.scale);
    TK_HIP(hipGetLastError());
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TOPK_OK;
}

}  // namespace topk

using namespace topk;

extern "C" {

int topk_learned_train(topk_ctx *ctx, uint32_t nq, const topk_query *qs, const topk_learned_config *cfg_in,
                       topk_learned_report *rep) {
    if (!ctx || !qs || nq < 16) return TOPK_ERR_INVALID;
    if (ctx->d > LRN_MAX_D) return TOPK_ERR_UNSUPPORTED;
    // real-data batches and stores with 64-bit ids run the FTE path: no threshold to learn
    if (ctx->score_div || ctx->n >= ID_LIMIT) return TOPK_ERR_UNSUPPORTED;
    topk_learned_config cfg;
    topk_learned_default_config(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (cfg.epochs < 1 || cfg.batch < 64 || cfg.batch % 64 || !(cfg.lr > 0) || !(cfg.target_fallback > 0) ||
        !(cfg.target_fallback < 1) || !(cfg.calib_fraction > 0) || !(cfg.calib_fraction < 1))
        return TOPK_ERR_INVALID;
    const double t0 = now_ms();
    LearnedState *Ls = lstate(ctx);
    int rc = build_sketch(ctx, Ls->sk);
    if (rc) return rc;
    Ls->sk_valid = true;

    // 1. labels: exact k-th scores from the batch path with seed thresholds
    std::vector<float> kth(nq);
    {
        const int saved = ctx->tune.batch_mode;
        ctx->tune.batch_mode = 1;
        const uint32_t chunk = 4096;
        std::vector<uint64_t> ids;
        std::vector<float> sc;
        for (uint32_t c0 = 0; c0 < nq && !rc; c0 += chunk) {
            const uint32_t c = nq - c0 < chunk ? nq - c0 : chunk;
            uint32_t ks = 1;
            for (uint32_t i = 0; i < c; i++) ks = qs[c0 + i].k > ks ? qs[c0 + i].k : ks;
            ids.resize((size_t)c * ks);
            sc.resize((size_t)c * ks);
            rc = topk_batch(ctx, c, qs + c0, ks, ids.data(), sc.data(), kth.data() + c0);
        }
        ctx->tune.batch_mode = saved;
        if (rc) return rc;
    }
    const double t1 = now_ms();

    // 2. features and targets
    MlpBuffers B;
    rc = features(ctx, Ls->sk, nq, qs, B);
    if (rc) return rc;
    std::vector<float> scale(nq), y(nq);
    TK_HIP(hipMemcpy(scale.data(), B.scale, sizeof(float) * nq, hipMemcpyDeviceToHost));
    std::vector<uint32_t> usable;
    for (uint32_t i = 0; i < nq; i++) {
        y[i] = scale[i] > 0 ? kth[i] / scale[i] : 0.0f;
        if (scale[i] > 0 && std::isfinite(y[i])) usable.push_back(i);
    }
    if (usable.size() < 16) return TOPK_ERR_INVALID;
    // saturation: labels at the query's largest score the sketch allows
    // (sum of normalised weight x sample maximum, feature 91)
    {
        uint32_t sat = 0;
        for (uint32_t i : usable) {
            float wa[LRN_MAX_D] = {0.0f};
            for (uint32_t j = 0; j < qs[i].qq; j++) wa[qs[i].attr[j]] += qs[i].weights[qs[i].attr[j]];
            float sabs = 0.0f, smax = 0.0f;
            for (int a = 0; a < LRN_MAX_D; a++) sabs += fabsf(wa[a]);
            for (int a = 0; a < LRN_MAX_D; a++) smax += wa[a] / sabs * Ls->sk.cmax[a];
            sat += y[i] >= smax - 1e-5f * fabsf(smax) ? 1u : 0u;
        }
        Ls->saturated = sat >= usable.size() / 200 + 1;  // at least 0.5% of the labels
    }
    std::mt19937_64 rng(cfg.seed);
    std::shuffle(usable.begin(), usable.end(), rng);
    uint32_t ncal = (uint32_t)(cfg.calib_fraction * usable.size());
    if (ncal < 8) ncal = 8;
    const uint32_t ntr = (uint32_t)usable.size() - ncal;
    if (ntr < 8) return TOPK_ERR_INVALID;
    std::vector<uint32_t> cal(usable.begin(), usable.begin() + ncal), tr(usable.begin() + ncal, usable.end());

    // 3. standardised copies of features and targets (training-set mean and
    // deviation; constant features map to 0), folded back into the first and
    // last layer after training so the installed network reads raw features
    std::vector<float> hx((size_t)nq * MLP_IN);
    TK_HIP(hipMemcpy(hx.data(), B.X, sizeof(float) * hx.size(), hipMemcpyDeviceToHost));
    double mu[MLP_IN], sd[MLP_IN], my = 0, sy = 0;
    for (int i = 0; i < MLP_IN; i++) {
        double a = 0, a2 = 0;
        for (uint32_t r : tr) {
            const double v = hx[(size_t)r * MLP_IN + i];
            a += v;
            a2 += v * v;
        }
        mu[i] = a / tr.size();
        const double var = a2 / tr.size() - mu[i] * mu[i];
        sd[i] = var > 1e-12 * (mu[i] * mu[i] + 1.0) ? std::sqrt(var) : 0.0;
    }
    for (uint32_t r : tr) my += y[r];
    my /= tr.size();
    for (uint32_t r : tr) sy += (y[r] - my) * (y[r] - my);
    sy = std::sqrt(sy / tr.size());
    if (!(sy > 0)) sy = 1.0;
    std::vector<float> hxs(hx.size()), ys(nq);
    for (uint32_t r = 0; r < nq; r++) {
        for (int i = 0; i < MLP_IN; i++)
            hxs[(size_t)r * MLP_IN + i] = sd[i] > 0 ? (float)((hx[(size_t)r * MLP_IN + i] - mu[i]) / sd[i]) : 0.0f;
        ys[r] = (float)((y[r] - my) / sy);
    }
    TK_HIP(hipMalloc(&B.Xs, sizeof(float) * hxs.size()));
    TK_HIP(hipMemcpy(B.Xs, hxs.data(), sizeof(float) * hxs.size(), hipMemcpyHostToDevice));

    // 4. parameters: uniform Glorot init
    std::vector<float> p(MLP_NP, 0.0f);
    {
        auto fill = [&](uint32_t off, uint32_t cnt, double lim) {
            std::uniform_real_distribution<double> u(-lim, lim);
            for (uint32_t i = 0; i < cnt; i++) p[off + i] = (float)u(rng);
        };
        fill(OFF_W1, MLP_H * MLP_IN, std::sqrt(6.0 / (MLP_IN + MLP_H)));
        fill(OFF_W2, MLP_H * MLP_H, std::sqrt(6.0 / (2 * MLP_H)));
        fill(OFF_W3, MLP_H, std::sqrt(6.0 / (MLP_H + 1)));
    }
    const uint32_t gmax = cfg.batch / MLP_ROWS;
    TK_HIP(hipMalloc(&B.Y, sizeof(float) * nq));
    TK_HIP(hipMalloc(&B.P, sizeof(float) * MLP_NP));
    TK_HIP(hipMalloc(&B.M, sizeof(float) * MLP_NP));
    TK_HIP(hipMalloc(&B.V, sizeof(float) * MLP_NP));
    TK_HIP(hipMalloc(&B.G, sizeof(float) * MLP_NP * gmax));
    TK_HIP(hipMalloc(&B.lsum, sizeof(float) * gmax));
    TK_HIP(hipMalloc(&B.rows, sizeof(uint32_t) * nq));
    TK_HIP(hipMalloc(&B.out, sizeof(float) * nq));
    TK_HIP(hipMemcpy(B.Y, ys.data(), sizeof(float) * nq, hipMemcpyHostToDevice));
    TK_HIP(hipMemcpy(B.P, p.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    TK_HIP(hipMemset(B.M, 0, sizeof(float) * MLP_NP));
    TK_HIP(hipMemset(B.V, 0, sizeof(float) * MLP_NP));
    rc = set_pass_lds();
    if (rc) return rc;

    // 5. mini-batch Adam on the standardised data
    hipStream_t s = ctx->stream;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double pb1 = 1.0, pb2 = 1.0;
    std::vector<uint32_t> perm = tr;
    for (uint32_t ep = 0; ep < cfg.epochs; ep++) {
        TK_HIP(hipStreamSynchronize(s));  // the previous epoch's copy of perm has been consumed
        std::shuffle(perm.begin(), perm.end(), rng);
        TK_HIP(hipMemcpyAsync(B.rows, perm.data(), sizeof(uint32_t) * ntr, hipMemcpyHostToDevice, s));
        const float lr = cfg.lr * (ep * 2 < cfg.epochs ? 1.0f : (ep * 4 < 3 * cfg.epochs ? 0.3f : 0.1f));
        for (uint32_t r0 = 0; r0 < ntr; r0 += cfg.batch) {
            const uint32_t nr = ntr - r0 < cfg.batch ? ntr - r0 : cfg.batch;
            if (nr < cfg.batch / 2 && r0 > 0) break;  // drop a short last batch
            rc = mlp_pass(s, B.P, B.Xs, B.Y, B.rows, r0, nr, 1.0f / (float)nr, nullptr, B.G, B.lsum, true);
            if (rc) return rc;
            pb1 *= b1;
            pb2 *= b2;
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, B.P, B.M, B.V, B.G,
                               (nr + MLP_ROWS - 1) / MLP_ROWS, lr, b1, b2, eps, (float)(1.0 - pb1),
                               (float)(1.0 - pb2), (float *)nullptr);
        }
        TK_HIP(hipGetLastError());
    }
    TK_HIP(hipStreamSynchronize(s));
    const double t2 = now_ms();

    // 6. fold the standardisation into W1, b1 (inputs) and W3, b3 (target)
    TK_HIP(hipMemcpy(p.data(), B.P, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    for (int j = 0; j < MLP_H; j++) {
        double b = p[OFF_B1 + j];
        for (int i = 0; i < MLP_IN; i++) {
            const double w = p[OFF_W1 + j * MLP_IN + i];
            if (sd[i] > 0) {
                p[OFF_W1 + j * MLP_IN + i] = (float)(w / sd[i]);
                b -= w * mu[i] / sd[i];
            } else {
                p[OFF_W1 + j * MLP_IN + i] = 0.0f;
            }
        }
        p[OFF_B1 + j] = (float)b;
    }
    for (int j = 0; j < MLP_H; j++) p[OFF_W3 + j] = (float)(p[OFF_W3 + j] * sy);
    p[OFF_B3] = (float)(p[OFF_B3] * sy + my);
    TK_HIP(hipMemcpy(B.P, p.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice));

    // 7. calibration of the installed network on held-out queries
    std::vector<uint32_t> all(tr);
    all.insert(all.end(), cal.begin(), cal.end());
    TK_HIP(hipMemcpy(B.rows, all.data(), sizeof(uint32_t) * all.size(), hipMemcpyHostToDevice));
    rc = mlp_pass(s, B.P, B.X, B.Y, B.rows, 0, (uint32_t)all.size(), 0.0f, B.out, nullptr, nullptr, false);
    if (rc) return rc;
    TK_HIP(hipStreamSynchronize(s));
    std::vector<float> pr(all.size());
    TK_HIP(hipMemcpy(pr.data(), B.out, sizeof(float) * all.size(), hipMemcpyDeviceToHost));
    double tl = 0, cl = 0;
    std::vector<float> resid(ncal);
    for (uint32_t i = 0; i < ntr; i++) tl += (double)(pr[i] - y[all[i]]) * (pr[i] - y[all[i]]);
    for (uint32_t i = 0; i < ncal; i++) {
        const float e = pr[ntr + i] - y[all[ntr + i]];
        cl += (double)e * e;
        resid[i] = e;
    }
    std::sort(resid.begin(), resid.end());
    size_t qi = (size_t)std::ceil((1.0 - cfg.target_fallback) * (double)ncal);
    if (qi >= ncal) qi = ncal - 1;
    float margin = resid[qi];
    margin = margin > 0 ? margin * 1.0001f + 1e-6f : 1e-6f;
    uint32_t over = 0;
    for (float r : resid) over += r > margin ? 1u : 0u;
    rc = install(ctx, B.P, margin);
    if (rc) return rc;
    if (rep) {
        memset(rep, 0, sizeof(*rep));
        rep->n_train = ntr;
        rep->n_calib = ncal;
        rep->train_mse = (float)(tl / ntr);
        rep->calib_mse = (float)(cl / ncal);
        rep->margin = margin;
        rep->calib_fallback = (float)over / (float)ncal;
        rep->ms_label = t1 - t0;
        rep->ms_train = t2 - t1;
        rep->saturated = Ls->saturated ? 1u : 0u;
    }
    return TOPK_OK;
}

int topk_learned_default_config(topk_learned_config *cfg) {
    if (!cfg) return TOPK_ERR_INVALID;
    cfg->epochs = 60;
    cfg->batch = 512;
    cfg->lr = 2e-3f;
    cfg->target_fallback = 0.0005f;  // a miss costs a second pass for the whole batch: keep misses rare
    cfg->calib_fraction = 0.2f;
    cfg->seed = 1;
    return TOPK_OK;
}

int topk_learned_clear(topk_ctx *ctx) {
    if (!ctx) return TOPK_ERR_INVALID;
    batch_quiesce(ctx);
    learned_free(ctx);
    return TOPK_OK;
}

int topk_learned_get(topk_ctx *ctx, float *params, float *margin) {
    if (!ctx) return TOPK_ERR_INVALID;
    if (!learned_ready(ctx)) return TOPK_ERR_STATE;
    LearnedState *L = lstate(ctx);
    if (params) TK_HIP(hipMemcpy(params, L->d_params, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    if (margin) *margin = L->margin;
    return TOPK_OK;
}

int topk_learned_set(topk_ctx *ctx, const float *params, float margin) {
    if (!ctx || !params || !std::isfinite(margin)) return TOPK_ERR_INVALID;
    if (ctx->d > LRN_MAX_D) return TOPK_ERR_UNSUPPORTED;
    LearnedState *L = lstate(ctx);
    if (!L->sk_valid) {
        const int rc = build_sketch(ctx, L->sk);
        if (rc) return rc;
        L->sk_valid = true;
    }
    return install(ctx, params, margin);
}

int topk_learned_features(topk_ctx *ctx, uint32_t nq, const topk_query *qs, float *x, float *scale) {
    if (!ctx || !qs || !nq) return TOPK_ERR_INVALID;
    if (ctx->d > LRN_MAX_D) return TOPK_ERR_UNSUPPORTED;
    LearnedState *L = lstate(ctx);
    if (!L->sk_valid) {
        const int rc = build_sketch(ctx, L->sk);
        if (rc) return rc;
        L->sk_valid = true;
    }
    MlpBuffers B;
    const int rc = features(ctx, L->sk, nq, qs, B);
    if (rc) return rc;
    if (x) TK_HIP(hipMemcpy(x, B.X, sizeof(float) * MLP_IN * nq, hipMemcpyDeviceToHost));
    if (scale) TK_HIP(hipMemcpy(scale, B.scale, sizeof(float) * nq, hipMemcpyDeviceToHost));
    return TOPK_OK;
}

int topk_learned_predict(topk_ctx *ctx, uint32_t nq, const topk_query *qs, float *pred, float *tau) {
    if (!ctx || !qs || !nq) return TOPK_ERR_INVALID;
    if (!learned_ready(ctx)) return TOPK_ERR_STATE;
    LearnedState *L = lstate(ctx);
    std::vector<BatchQuery> hq(nq);
    for (uint32_t i = 0; i < nq; i++) {
        const int rc = make_batch_query(ctx, qs[i], hq[i]);
        if (rc) return rc;
    }
    MlpBuffers B;
    TK_HIP(hipMalloc(&B.q, sizeof(BatchQuery) * nq));
    TK_HIP(hipMalloc(&B.out, sizeof(float) * nq));
    TK_HIP(hipMemcpy(B.q, hq.data(), sizeof(BatchQuery) * nq, hipMemcpyHostToDevice));
    uint32_t g = (nq + TAU_BS / 64 - 1) / (TAU_BS / 64);
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(mlp_tau_kernel, dim3(g), dim3(TAU_BS), 0, ctx->stream, L->d_params, B.q, nq, L->sk, L->margin,
                       B.out, 3.0f, 0);
    TK_HIP(hipGetLastError());
    TK_HIP(hipStreamSynchronize(ctx->stream));
    if (pred) TK_HIP(hipMemcpy(pred, B.out, sizeof(float) * nq, hipMemcpyDeviceToHost));
    if (tau) {
        TK_HIP(hipMemcpy(hq.data(), B.q, sizeof(BatchQuery) * nq, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nq; i++) tau[i] = hq[i].tau;
    }
    return TOPK_OK;
}

int topk_mlp_eval(topk_ctx *ctx, const float *params, const float *x, const float *y, uint32_t nb, float *out,
                  float *grad, float *loss) {
    if (!ctx || !params || !x || !nb || ((grad || loss) && !y)) return TOPK_ERR_INVALID;
    if (nb > 65536) return TOPK_ERR_UNSUPPORTED;
    MlpBuffers B;
    const uint32_t g = (nb + MLP_ROWS - 1) / MLP_ROWS;
    TK_HIP(hipMalloc(&B.P, sizeof(float) * MLP_NP));
    TK_HIP(hipMalloc(&B.X, sizeof(float) * MLP_IN * nb));
    TK_HIP(hipMalloc(&B.Y, sizeof(float) * nb));
    TK_HIP(hipMalloc(&B.out, sizeof(float) * nb));
    TK_HIP(hipMalloc(&B.G, sizeof(float) * MLP_NP * g));
    TK_HIP(hipMalloc(&B.M, sizeof(float) * MLP_NP));
    TK_HIP(hipMalloc(&B.lsum, sizeof(float) * g));
    TK_HIP(hipMemcpy(B.P, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    TK_HIP(hipMemcpy(B.X, x, sizeof(float) * MLP_IN * nb, hipMemcpyHostToDevice));
    if (y) TK_HIP(hipMemcpy(B.Y, y, sizeof(float) * nb, hipMemcpyHostToDevice));
    int rc = set_pass_lds();
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const bool bw = grad || loss;
    rc = mlp_pass(s, B.P, B.X, B.Y, nullptr, 0, nb, 1.0f / (float)nb, B.out, bw ? B.G : nullptr,
                  bw ? B.lsum : nullptr, bw);
    if (rc) return rc;
    if (bw)
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, B.P, B.M, B.M, B.G, g, 0.0f,
                           0.0f, 0.0f, 0.0f, 1.0f, 1.0f, B.M);
    TK_HIP(hipGetLastError());
    TK_HIP(hipStreamSynchronize(s));
    if (out) TK_HIP(hipMemcpy(out, B.out, sizeof(float) * nb, hipMemcpyDeviceToHost));
    if (grad) TK_HIP(hipMemcpy(grad, B.M, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    if (loss) {
        std::vector<float> l(g);
        TK_HIP(hipMemcpy(l.data(), B.lsum, sizeof(float) * g, hipMemcpyDeviceToHost));
        double t = 0;
        for (float v : l) t += v;
        *loss = (float)(t / nb);
    }
    return TOPK_OK;
}

}  // extern "C"
