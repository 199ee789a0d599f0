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
//    hand-written forward / backward / update kernels (mlp_pass_kernel on
//    FP32 matrix cores, mlp_adam_kernel); gradients match the FP64 oracle
//    restatement.
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

#include <hip/hip_cooperative_groups.h>

#include "device.hpp"
#include "internal.hpp"

namespace topk {

constexpr int MLP_IN = 96, MLP_H = 64;
constexpr uint32_t OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H,
                   OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H,
                   MLP_NP = OFF_B3 + 1;
static_assert(MLP_NP == 10433, "parameter count");
constexpr int LRN_MAX_D = 16;
constexpr int MLP_BS = 256;
constexpr int MLP_TR = 16;     // samples per forward / backward workgroup
constexpr uint32_t MLP_MAX_BATCH = 4096;       // training mini-batch cap: batch / 16 co-resident workgroups
constexpr uint32_t MLP_EPOCH_ROWS = 1u << 24;  // row-order entries per training launch

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
    // every term slot's weight and attribute read at once (16-byte loads; a
    // loop bounded by m would wait on memory once per term), then the
    // weights accumulated per attribute in term order
    const uint32_t m = Q.m;
    float wt[TOPK_MAX_DIMS];
    uint32_t ap[TOPK_MAX_DIMS / 4];  // four attribute bytes per word
#pragma unroll
    for (int i = 0; i < TOPK_MAX_DIMS; i += 4) {
        const f4 v = *reinterpret_cast<const f4 *>(&Q.w[i]);
        wt[i] = v[0];
        wt[i + 1] = v[1];
        wt[i + 2] = v[2];
        wt[i + 3] = v[3];
    }
#pragma unroll
    for (int i = 0; i < TOPK_MAX_DIMS / 16; i++) {
        const uint4 v = *reinterpret_cast<const uint4 *>(&Q.attr[16 * i]);
        ap[4 * i] = v.x;
        ap[4 * i + 1] = v.y;
        ap[4 * i + 2] = v.z;
        ap[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < TOPK_MAX_DIMS; j++) {
        const uint32_t at = (ap[j / 4] >> (8 * (j % 4))) & 0xFFu;
        const bool on = (uint32_t)j < m;
#pragma unroll
        for (int a = 0; a < LRN_MAX_D; a++) wa[a] = on && (uint32_t)a == at ? wa[a] + wt[j] : wa[a];
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
// forward / backward over a tile of 16 samples on FP32 matrix cores
// ---------------------------------------------------------------------------
// v_mfma_f32_16x16x4_f32 computes D[16][16] += A[16][4] * B[4][16] across a
// wave: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives
// D[4 * (l / 16) + r][l % 16] for r = 0..3.  A product over K inputs is K / 4
// such steps; the four K slots of a step may stand for any four inputs, so
// lane group g = l / 16 takes a contiguous run of inputs (g * K / 4 + s at
// step s) and reads both operands as 16-byte rows.
//
// Workgroup: 16 samples, 4 waves; wave w owns hidden units [16w, 16w + 16)
// of both layers (the N of every forward product), so a lane's D entries
// are (sample 4g + r, unit 16w + l % 16).  The backward products reuse that
// placement: dW1 and dW2 for wave w's units take the sample index as K with
// the same 4g + s split, which makes A = the lane's own G1 / G2 registers.
constexpr int MLP_XP = MLP_IN + 4, MLP_HP = MLP_H + 4;  // LDS row pitches (bank spread)

struct MlpTile {
    float X[MLP_TR][MLP_XP];
    float H1[MLP_TR][MLP_HP];
    float G2[MLP_TR][MLP_HP];
    float part[4][MLP_TR];                     // per-wave output partial sums
    float go[MLP_TR], e2[MLP_TR];
};

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// v rotated by N lanes within each 16-lane row (DPP row_ror:N).
template <int N>
__device__ __forceinline__ float row_ror(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + N, 0xF, 0xF, false));
}

// v[l] + v[l ^ 16] + v[l ^ 32] + v[l ^ 48] in every lane l (the four lanes
// that share l % 16), through the lane-swap instructions.
__device__ __forceinline__ float cross_row_sum4(float v) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Sum over the 16 lanes of v's DPP row, in every lane of the row.
__device__ __forceinline__ float row_sum16(float v) {
    v += row_ror<1>(v);
    v += row_ror<2>(v);
    v += row_ror<4>(v);
    v += row_ror<8>(v);
    return v;
}

// One tile: forward of samples rows[base + b] (INDIRECT; else base + b)
// for 1 <= b < nr <= 16 and, when BACKWARD, the gradient of
// sum_b (out_b - y_b)^2 * inv_total into gp[0 .. MLP_NP) and the squared
// errors' sum into *lsum.  Called by all 256 threads of the workgroup.
template <bool BACKWARD, bool INDIRECT>
__device__ __forceinline__ void mlp_tile(MlpTile &T, const float *P, const float *__restrict__ X,
                                         const float *__restrict__ Y, const uint32_t *__restrict__ rows,
                                         uint32_t base, uint32_t nr, float inv_total, float *__restrict__ out,
                                         float *__restrict__ gp, float *__restrict__ lsum) {
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / 64;
    const uint32_t c = lane & 15, g = lane >> 4;   // D column / A row, K slot group
    const uint32_t unit = 16 * w + c;              // hidden unit this lane's B column stands for
    // Every load of the tile is issued unconditionally (rows past nr read
    // sample nr - 1 and are zeroed afterwards) so none waits on a branch.
    // B operands: W1[unit][24g .. 24g + 24), W2[unit][16g .. 16g + 16)
    f4 w1[6], w2[4];
    {
        const f4 *p1 = reinterpret_cast<const f4 *>(P + OFF_W1 + unit * MLP_IN + 24 * g);
        const f4 *p2 = reinterpret_cast<const f4 *>(P + OFF_W2 + unit * MLP_H + 16 * g);
#pragma unroll
        for (int t = 0; t < 6; t++) w1[t] = p1[t];
#pragma unroll
        for (int t = 0; t < 4; t++) w2[t] = p2[t];
    }
    const float b1 = P[OFF_B1 + unit], b2 = P[OFF_B2 + unit], w3 = P[OFF_W3 + unit], b3 = P[OFF_B3];
    // backward B operand of G1 = G2 W2: W2[16g + s][unit], a column slice
    float w2c[16];
    if constexpr (BACKWARD) {
#pragma unroll
        for (int t = 0; t < 16; t++) w2c[t] = P[OFF_W2 + (16 * g + t) * MLP_H + unit];
    }
    // sample tile (16 x 96 as 384 float4): thread tid covers slots tid and tid + 256
    constexpr uint32_t X4 = MLP_IN / 4, NX4 = MLP_TR * X4;
    // (threads past slot 383 repeat their first slot: same value, same place)
    const uint32_t oA = tid, oB = tid + MLP_BS < NX4 ? tid + MLP_BS : tid;
    const uint32_t bA = oA / X4, bB = oB / X4;
    const uint32_t sA = bA < nr ? bA : nr - 1, sB = bB < nr ? bB : nr - 1, sY = tid < nr ? tid : nr - 1;
    const uint32_t rA = INDIRECT ? rows[base + sA] : base + sA, rB = INDIRECT ? rows[base + sB] : base + sB;
    float yv = 0.0f;  // target of sample tid
    if constexpr (BACKWARD) yv = Y[INDIRECT ? rows[base + sY] : base + sY];
    const f4 xA = reinterpret_cast<const f4 *>(X + (uint64_t)rA * MLP_IN)[oA % X4];
    const f4 xB = reinterpret_cast<const f4 *>(X + (uint64_t)rB * MLP_IN)[oB % X4];
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    *reinterpret_cast<f4 *>(&T.X[bA][4 * (oA % X4)]) = bA < nr ? xA : zero;
    *reinterpret_cast<f4 *>(&T.X[bB][4 * (oB % X4)]) = bB < nr ? xB : zero;
    __syncthreads();
    // Operands of the backward products are read into registers as soon as
    // their producer's barrier has passed, so those products wait on no LDS.

    // layer 1: H1[b][unit] = relu(b1 + sum_i X[b][i] W1[unit][i]), K = 96
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const f4 a = *reinterpret_cast<const f4 *>(&T.X[c][24 * g + 4 * t]);
#pragma unroll
        for (int e = 0; e < 4; e++) acc = mfma4(a[e], w1[t][e], acc);
    }
    float xb[MLP_IN / 16][4];  // dW1's B operand: X[4g + s][16 it + c]
    if constexpr (BACKWARD) {
#pragma unroll
        for (int it = 0; it < MLP_IN / 16; it++)
#pragma unroll
            for (int t = 0; t < 4; t++) xb[it][t] = T.X[4 * g + t][16 * it + c];
    }
    float h1[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float s = acc[r] + b1;
        h1[r] = s > 0.0f ? s : 0.0f;
        T.H1[4 * g + r][unit] = h1[r];
    }
    __syncthreads();

    // layer 2: H2[b][unit] = relu(b2 + sum_i H1[b][i] W2[unit][i]), K = 64
    acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const f4 a = *reinterpret_cast<const f4 *>(&T.H1[c][16 * g + 4 * t]);
#pragma unroll
        for (int e = 0; e < 4; e++) acc = mfma4(a[e], w2[t][e], acc);
    }
    float hb[MLP_H / 16][4];  // dW2's B operand: H1[4g + s][16 it + c]
    if constexpr (BACKWARD) {
#pragma unroll
        for (int it = 0; it < MLP_H / 16; it++)
#pragma unroll
            for (int t = 0; t < 4; t++) hb[it][t] = T.H1[4 * g + t][16 * it + c];
    }
    float h2[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float s = acc[r] + b2;
        h2[r] = s > 0.0f ? s : 0.0f;
        // output: sum over this wave's 16 units (lanes sharing g), then over waves
        const float p = row_sum16(w3 * h2[r]);  // lanes sharing g form one DPP row
        if (c == 0) T.part[w][4 * g + r] = p;
    }
    __syncthreads();
    if (tid < MLP_TR) {
        const float y = ((T.part[0][tid] + T.part[1][tid]) + (T.part[2][tid] + T.part[3][tid])) + b3;
        float e = 0.0f;
        if (tid < nr) {
            if (out) out[base + tid] = y;
            if constexpr (BACKWARD) e = y - yv;
        }
        T.go[tid] = 2.0f * e * inv_total;
        T.e2[tid] = e * e;
    }
    if constexpr (!BACKWARD) return;
    __syncthreads();

    // G2 = dL/d(pre-activation 2); dW3, db2 per unit (sums over the 16 samples)
    float g2[4], sw3 = 0.0f, sb2 = 0.0f;
    const f4 go4 = *reinterpret_cast<const f4 *>(&T.go[4 * g]);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        g2[r] = h2[r] > 0.0f ? go4[r] * w3 : 0.0f;
        T.G2[4 * g + r][unit] = g2[r];
        sw3 += go4[r] * h2[r];
        sb2 += g2[r];
    }
    sw3 = cross_row_sum4(sw3);
    sb2 = cross_row_sum4(sb2);
    if (g == 0) {
        gp[OFF_W3 + unit] = sw3;
        gp[OFF_B2 + unit] = sb2;
    }
    // dW2[unit rows of wave w][i] = sum_b G2[b][j] H1[b][i]: K = samples, A = g2
#pragma unroll
    for (int it = 0; it < MLP_H / 16; it++) {
        f4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 4; s++) d = mfma4(g2[s], hb[it][s], d);
#pragma unroll
        for (int r = 0; r < 4; r++) gp[OFF_W2 + (16 * w + 4 * g + r) * MLP_H + 16 * it + c] = d[r];
    }
    __syncthreads();

    // G1[b][unit] = [H1 > 0] sum_j G2[b][j] W2[j][unit], K = 64
    acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const f4 a = *reinterpret_cast<const f4 *>(&T.G2[c][16 * g + 4 * t]);
#pragma unroll
        for (int e = 0; e < 4; e++) acc = mfma4(a[e], w2c[4 * t + e], acc);
    }
    float g1[4], sb1 = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        g1[r] = h1[r] > 0.0f ? acc[r] : 0.0f;
        sb1 += g1[r];
    }
    sb1 = cross_row_sum4(sb1);
    if (g == 0) gp[OFF_B1 + unit] = sb1;
    // dW1[unit rows of wave w][i] = sum_b G1[b][j] X[b][i]: K = samples, A = g1
#pragma unroll
    for (int it = 0; it < MLP_IN / 16; it++) {
        f4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 4; s++) d = mfma4(g1[s], xb[it][s], d);
#pragma unroll
        for (int r = 0; r < 4; r++) gp[OFF_W1 + (16 * w + 4 * g + r) * MLP_IN + 16 * it + c] = d[r];
    }
    if (tid == 0) {
        float s = 0.0f, l = 0.0f;
        for (int b = 0; b < MLP_TR; b++) {
            s += T.go[b];
            l += T.e2[b];
        }
        gp[OFF_B3] = s;
        *lsum = l;
    }
}

// Forward of rows rows[r0 + b] (b < nrows - r0; rows == nullptr: r0 + b), 16
// per workgroup and, when BACKWARD, workgroup i's partial gradient into
// G[i * MLP_NP ...] and its squared errors' sum into lsum[i].
template <bool BACKWARD>
__global__ __launch_bounds__(MLP_BS) void mlp_pass_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                          const float *__restrict__ Y,
                                                          const uint32_t *__restrict__ rows, uint32_t r0,
                                                          uint32_t nrows, float inv_total, float *__restrict__ out,
                                                          float *__restrict__ G, float *__restrict__ lsum) {
    __shared__ __align__(16) MlpTile T;
    const uint32_t base = r0 + blockIdx.x * MLP_TR;
    const uint32_t nr = base < nrows ? (nrows - base < (uint32_t)MLP_TR ? nrows - base : (uint32_t)MLP_TR) : 0u;
    if (rows)
        mlp_tile<BACKWARD, true>(T, P, X, Y, rows, base, nr, inv_total, out,
                                 BACKWARD ? G + (uint64_t)blockIdx.x * MLP_NP : G, BACKWARD ? lsum + blockIdx.x : lsum);
    else
        mlp_tile<BACKWARD, false>(T, P, X, Y, rows, base, nr, inv_total, out,
                                  BACKWARD ? G + (uint64_t)blockIdx.x * MLP_NP : G, BACKWARD ? lsum + blockIdx.x : lsum);
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

// Xp[i] = X[perm[i]] and Yp[i] = Y[perm[i]] for i < n: one epoch's samples
// in its row order, so a training step reads contiguous rows.
__global__ void mlp_gather_kernel(const float *__restrict__ X, const float *__restrict__ Y,
                                  const uint32_t *__restrict__ perm, uint32_t n, float *__restrict__ Xp,
                                  float *__restrict__ Yp) {
    constexpr uint32_t X4 = MLP_IN / 4;
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < n * X4; o += gridDim.x * blockDim.x) {
        const uint32_t i = o / X4, r = perm[i];
        reinterpret_cast<f4 *>(Xp)[o] = reinterpret_cast<const f4 *>(X)[(uint64_t)r * X4 + o % X4];
        if (o % X4 == 0) Yp[i] = Y[r];
    }
}

// One optimisation step of the training kernel: step size, Adam bias
// corrections 1 - b1^t and 1 - b2^t, 1 / rows in the step.
struct MlpStep {
    float lr, bc1, bc2, inv_rows;
};

struct MlpTrainArgs {
    float *P, *M, *V;        // parameters and Adam moments, updated in place
    float *G;                // gridDim.x x MLP_NP partial gradients
    float *lsum;             // gridDim.x squared-error sums
    const float *X, *Y;      // standardised samples and targets
    const uint32_t *perm;    // epochs x ntr: each epoch's order of the training rows
    const MlpStep *sched;    // epochs x steps_per_epoch
    uint32_t ntr, batch, steps_per_epoch, epochs;
    float b1, b2, eps;
};

constexpr uint32_t MLP_NCOLS = (MLP_NP + 31) / 32;  // 32-parameter columns

// Sums the nb partial gradients of the 32-parameter columns c0 + u * cstride
// (u < NC) and applies Adam to them.  Thread group j = tid / 32 adds the
// partials b = j mod 8 (all loads of a round in flight; out-of-range ones
// read partial 0 and are dropped), then the 8 group sums are added in order,
// so the result does not depend on the launch shape.  All 256 threads call it.
template <int NC>
__device__ __forceinline__ void mlp_reduce_adam(uint32_t c0, uint32_t cstride, uint32_t nb, float *P, float *M,
                                                float *V, const float *G, const MlpStep &st, float b1, float b2,
                                                float eps, float (*red)[NC * 32]) {
    const uint32_t tid = threadIdx.x, q = tid & 31, j = tid >> 5;
    // this thread's parameters and moments, loaded alongside the partials
    constexpr int NS = (NC * 32 + MLP_BS - 1) / MLP_BS;
    float pp[NS], pm[NS], pv[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const uint32_t sl = tid + k * MLP_BS, p = (c0 + (sl / 32) * cstride) * 32 + sl % 32;
        const uint32_t pc = sl < NC * 32 && p < MLP_NP ? p : 0;
        pp[k] = P[pc];
        pm[k] = M[pc];
        pv[k] = V[pc];
    }
    float v[NC][4];
#pragma unroll
    for (int u = 0; u < NC; u++)
#pragma unroll
        for (int t = 0; t < 4; t++) v[u][t] = 0.0f;
    for (uint32_t b0 = j; b0 < nb; b0 += 32) {
        float x[NC][4];
#pragma unroll
        for (int u = 0; u < NC; u++) {
            const uint32_t p = (c0 + u * cstride) * 32 + q;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t b = b0 + 8 * t;
                x[u][t] = G[p < MLP_NP && b < nb ? (uint64_t)b * MLP_NP + p : 0];
            }
        }
#pragma unroll
        for (int u = 0; u < NC; u++) {
            const uint32_t p = (c0 + u * cstride) * 32 + q;
#pragma unroll
            for (int t = 0; t < 4; t++) v[u][t] += p < MLP_NP && b0 + 8 * t < nb ? x[u][t] : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < NC; u++) red[j][u * 32 + q] = (v[u][0] + v[u][1]) + (v[u][2] + v[u][3]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const uint32_t sl = tid + k * MLP_BS, p = (c0 + (sl / 32) * cstride) * 32 + sl % 32;
        if (sl >= NC * 32 || p >= MLP_NP) continue;
        float g = 0.0f;
#pragma unroll
        for (int jj = 0; jj < 8; jj++) g += red[jj][sl];
        const float m = b1 * pm[k] + (1.0f - b1) * g;
        const float vv = b2 * pv[k] + (1.0f - b2) * g * g;
        M[p] = m;
        V[p] = vv;
        const float mh = m / st.bc1, vh = vv / st.bc2;
        P[p] = pp[k] - st.lr * mh / (sqrtf(vh) + eps);
    }
    __syncthreads();
}

// The Adam half of one training step as its own launch: workgroup i takes
// columns [MLP_STEP_COLS i, MLP_STEP_COLS (i + 1)).
constexpr int MLP_STEP_COLS = 4;
__global__ __launch_bounds__(MLP_BS) void mlp_step_adam_kernel(float *__restrict__ P, float *__restrict__ M,
                                                               float *__restrict__ V, const float *__restrict__ G,
                                                               uint32_t nb, MlpStep st, float b1, float b2,
                                                               float eps) {
    __shared__ float red[8][MLP_STEP_COLS * 32];
    mlp_reduce_adam<MLP_STEP_COLS>(blockIdx.x * MLP_STEP_COLS, 1, nb, P, M, V, G, st, b1, b2, eps, red);
}

// Every mini-batch step of `epochs` epochs in one cooperative launch of
// batch / 16 co-resident workgroups.  Step: workgroup i runs the tile of
// rows [16 i, 16 i + 16) of the batch into its partial gradient; grid
// barrier; workgroup i sums the partials of the columns i, i + grid, ... and
// applies Adam to them (mlp_reduce_adam); grid barrier.  The moments of a
// column are only touched by its owner; parameters and partials cross
// workgroups through the barriers.
constexpr int MLP_RED_COLS = 12;  // columns a workgroup reduces per round
__global__ __launch_bounds__(MLP_BS) void mlp_train_kernel(MlpTrainArgs a) {
    __shared__ __align__(16) MlpTile T;
    __shared__ float red[8][MLP_RED_COLS * 32];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    for (uint32_t ep = 0; ep < a.epochs; ep++) {
        const uint32_t *rows = a.perm + (uint64_t)ep * a.ntr;
        for (uint32_t i = 0; i < a.steps_per_epoch; i++) {
            const MlpStep st = a.sched[ep * a.steps_per_epoch + i];
            const uint32_t r0 = i * a.batch, nrs = a.ntr - r0 < a.batch ? a.ntr - r0 : a.batch;
            const uint32_t nb = (nrs + MLP_TR - 1) / MLP_TR;
            if (blockIdx.x < nb) {
                const uint32_t base = r0 + blockIdx.x * MLP_TR;
                const uint32_t nr = r0 + nrs - base < (uint32_t)MLP_TR ? r0 + nrs - base : (uint32_t)MLP_TR;
                mlp_tile<true, true>(T, a.P, a.X, a.Y, rows, base, nr, st.inv_rows, nullptr,
                               a.G + (uint64_t)blockIdx.x * MLP_NP, a.lsum + blockIdx.x);
            }
            grid.sync();
            for (uint32_t c0 = blockIdx.x; c0 < MLP_NCOLS; c0 += gridDim.x * MLP_RED_COLS)
                mlp_reduce_adam<MLP_RED_COLS>(c0, gridDim.x, nb, a.P, a.M, a.V, a.G, st, a.b1, a.b2, a.eps, red);
            grid.sync();
        }
    }
}

// Threshold per batch query: tau = (prediction - margin) * sum |W| (or
// -inf when the weights are all zero).  pred (optional) gets the unlowered
// prediction in score units.  One workgroup per 16 queries: lanes 0..15 of
// the first wave build the queries' feature rows in the LDS tile while
// every wave loads its B operands, then the forward pass runs on FP32
// matrix cores with exactly the operation sequence of mlp_tile's forward
// part -- the predictions equal those the calibration computed (mlp_pass
// over the same features), bit for bit, and the weights are read once per
// 16 queries instead of once per query.
__global__ __launch_bounds__(MLP_BS) void mlp_tau_kernel(const float *__restrict__ P, BatchQuery *__restrict__ qs,
                                                         uint32_t nq, Sketch S, float margin,
                                                         float *__restrict__ pred, float retry_width,
                                                         int saturated) {
    __shared__ MlpTile T;
    __shared__ float fsc[MLP_TR], fmx[MLP_TR];  // per query: sum |W| and feature 91 (largest score / sum |W|)
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / 64;
    const uint32_t c = lane & 15, g = lane >> 4;
    const uint32_t unit = 16 * w + c;
    const uint32_t base = blockIdx.x * MLP_TR;
    const uint32_t nr = nq - base < MLP_TR ? nq - base : MLP_TR;
    f4 w1[6], w2[4];
    {
        const f4 *p1 = reinterpret_cast<const f4 *>(P + OFF_W1 + unit * MLP_IN + 24 * g);
        const f4 *p2 = reinterpret_cast<const f4 *>(P + OFF_W2 + unit * MLP_H + 16 * g);
#pragma unroll
        for (int t = 0; t < 6; t++) w1[t] = p1[t];
#pragma unroll
        for (int t = 0; t < 4; t++) w2[t] = p2[t];
    }
    const float b1 = P[OFF_B1 + unit], b2 = P[OFF_B2 + unit], w3 = P[OFF_W3 + unit], b3 = P[OFF_B3];
    if (tid < MLP_TR) {
        float f[MLP_IN], sc = 0.0f;
        if (tid < nr) {
            mlp_features(qs[base + tid], S, f, &sc);
        } else {
#pragma unroll
            for (int i = 0; i < MLP_IN; i++) f[i] = 0.0f;  // rows past nr: zero, as mlp_tile pads them
        }
#pragma unroll
        for (int i = 0; i < MLP_IN; i += 4) *reinterpret_cast<f4 *>(&T.X[tid][i]) = f4{f[i], f[i + 1], f[i + 2], f[i + 3]};
        fsc[tid] = sc;
        fmx[tid] = f[91];
    }
    __syncthreads();
    // layer 1 (K = 96), as in mlp_tile
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const f4 a = *reinterpret_cast<const f4 *>(&T.X[c][24 * g + 4 * t]);
#pragma unroll
        for (int e = 0; e < 4; e++) acc = mfma4(a[e], w1[t][e], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float v = acc[r] + b1;
        T.H1[4 * g + r][unit] = v > 0.0f ? v : 0.0f;
    }
    __syncthreads();
    // layer 2 (K = 64) and the output's partial sums, as in mlp_tile
    acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const f4 a = *reinterpret_cast<const f4 *>(&T.H1[c][16 * g + 4 * t]);
#pragma unroll
        for (int e = 0; e < 4; e++) acc = mfma4(a[e], w2[t][e], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float v = acc[r] + b2;
        const float h2 = v > 0.0f ? v : 0.0f;
        const float p = row_sum16(w3 * h2);
        if (c == 0) T.part[w][4 * g + r] = p;
    }
    __syncthreads();
    if (tid < nr) {
        const uint32_t q = base + tid;
        const float y = ((T.part[0][tid] + T.part[1][tid]) + (T.part[2][tid] + T.part[3][tid])) + b3;
        const float sc = fsc[tid];
        qs[q].tau = sc > 0.0f ? (y - margin) * sc : -INFINITY;
        // a miss reruns retry_width margins below the prediction, or
        // (retry_width < 0) with a seed threshold
        const float wide = retry_width * fmaxf(fabsf(margin), 1e-3f);  // strictly below tau
        qs[q].tau_retry = retry_width < 0.0f ? INFINITY : (sc > 0.0f ? (y - wide) * sc : -INFINITY);
        // within two margins of the query's largest score (feature 91):
        // ties likely, seed this query
        qs[q].seed = saturated && sc > 0.0f && y + 2.0f * fabsf(margin) >= fmx[tid] ? 1u : 0u;
        if (pred) pred[q] = y * sc;
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


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
    uint32_t *rows = nullptr, *perm = nullptr;
    float *Xp = nullptr, *Yp = nullptr;
    void *sched = nullptr;
    BatchQuery *q = nullptr;
    ~MlpBuffers() {
        hipFree(perm);
        hipFree(Xp);
        hipFree(Yp);
        hipFree(sched);
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

// Forward (+ backward) of nrows rows [r0, r0 + nrows) of X (through rows[]
// when given): one workgroup per 16 rows; workgroup b's partial gradient ->
// G + b * MLP_NP and its squared-error sum -> lsum[b].
static int mlp_pass(hipStream_t s, const float *P, const float *X, const float *Y, const uint32_t *rows, uint32_t r0,
                    uint32_t nrows, float inv_total, float *out, float *G, float *lsum, bool backward) {
    const uint32_t nb = (nrows + MLP_TR - 1) / MLP_TR;
    if (!nb) return TOPK_OK;
    if (backward)
        hipLaunchKernelGGL(mlp_pass_kernel<true>, dim3(nb), dim3(MLP_BS), 0, s, P, X, Y, rows, r0, r0 + nrows,
                           inv_total, out, G, lsum);
    else
        hipLaunchKernelGGL(mlp_pass_kernel<false>, dim3(nb), dim3(MLP_BS), 0, s, P, X, Y, rows, r0,
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
    const uint32_t g = (nq + MLP_TR - 1) / MLP_TR;
    const float width =
        ctx->tune.batch_retry == 2 ? -1.0f : (ctx->tune.batch_retry_width ? (float)ctx->tune.batch_retry_width : 3.0f);
    hipLaunchKernelGGL(mlp_tau_kernel, dim3(g), dim3(MLP_BS), 0, s, L->d_params, q, nq, L->sk, L->margin,
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
    if (cfg.epochs < 1 || cfg.batch < 64 || cfg.batch % 64 || cfg.batch > MLP_MAX_BATCH || !(cfg.lr > 0) || !(cfg.target_fallback > 0) ||
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
    if (Ls->saturated) {  // saturated queries find their threshold in the extreme-value bitmaps
        const int brc = pta_extreme_bits(ctx);
        if (brc) return brc;
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
    const uint32_t gmax = cfg.batch / MLP_TR;
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

    // 5. mini-batch Adam on the standardised data.  The epochs' row orders
    //    and per-step scalars are laid out on the host for up to
    //    MLP_EPOCH_ROWS rows' worth of epochs at a time; then either each
    //    epoch gathers its rows (mlp_gather_kernel) and each step is two
    //    launches, the MFMA pass and the reduce + Adam (default), or one
    //    cooperative launch runs every step (mlp_train_kernel, tuning
    //    "mlp_train" 1; slower here: its grid barriers cost more than a
    //    launch boundary).  Both give the same parameters.
    hipStream_t s = ctx->stream;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double pb1 = 1.0, pb2 = 1.0;
    std::vector<uint32_t> steps_rows;  // rows in each step of an epoch (a short last batch is dropped)
    for (uint32_t r0 = 0; r0 < ntr; r0 += cfg.batch) {
        const uint32_t nr = ntr - r0 < cfg.batch ? ntr - r0 : cfg.batch;
        if (nr < cfg.batch / 2 && r0 > 0) break;
        steps_rows.push_back(nr);
    }
    const uint32_t spe = (uint32_t)steps_rows.size();
    const uint32_t ep_chunk = std::max<uint32_t>(1, std::min<uint32_t>(cfg.epochs, MLP_EPOCH_ROWS / ntr));
    std::vector<uint32_t> perm = tr, perms((size_t)ep_chunk * ntr);
    std::vector<MlpStep> sched((size_t)ep_chunk * spe);
    TK_HIP(hipMalloc(&B.perm, sizeof(uint32_t) * perms.size()));
    TK_HIP(hipMalloc(&B.Xp, sizeof(float) * MLP_IN * ntr));
    TK_HIP(hipMalloc(&B.Yp, sizeof(float) * ntr));
    TK_HIP(hipMalloc(&B.sched, sizeof(MlpStep) * sched.size()));
    MlpStep *d_sched = static_cast<MlpStep *>(B.sched);
    MlpTrainArgs ta{B.P, B.M, B.V, B.G, B.lsum, B.Xs, B.Y, B.perm, d_sched, ntr, cfg.batch, spe, 0, b1, b2, eps};
    hipEvent_t ev[2] = {nullptr, nullptr};
    TK_HIP(hipEventCreate(&ev[0]));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 2; i++)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } ev_guard{ev};
    TK_HIP(hipEventCreate(&ev[1]));
    double ms_steps = 0.0, flops = 0.0;
    uint32_t nsteps = 0;
    for (uint32_t ep0 = 0; ep0 < cfg.epochs; ep0 += ep_chunk) {
        const uint32_t ne = std::min(ep_chunk, cfg.epochs - ep0);
        for (uint32_t e = 0; e < ne; e++) {
            const uint32_t ep = ep0 + e;
            std::shuffle(perm.begin(), perm.end(), rng);
            std::copy(perm.begin(), perm.end(), perms.begin() + (size_t)e * ntr);
            const float lr = cfg.lr * (ep * 2 < cfg.epochs ? 1.0f : (ep * 4 < 3 * cfg.epochs ? 0.3f : 0.1f));
            for (uint32_t i = 0; i < spe; i++) {
                pb1 *= b1;
                pb2 *= b2;
                sched[(size_t)e * spe + i] = MlpStep{lr, (float)(1.0 - pb1), (float)(1.0 - pb2),
                                                     1.0f / (float)steps_rows[i]};
            }
        }
        // synchronous copies: the previous launch has finished with both tables
        TK_HIP(hipMemcpy(B.perm, perms.data(), sizeof(uint32_t) * ne * ntr, hipMemcpyHostToDevice));
        TK_HIP(hipMemcpy(d_sched, sched.data(), sizeof(MlpStep) * ne * spe, hipMemcpyHostToDevice));
        TK_HIP(hipEventRecord(ev[0], s));
        if (ctx->tune.mlp_train == 1) {
            ta.epochs = ne;
            void *args[] = {&ta};
            TK_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(mlp_train_kernel),
                                              dim3(cfg.batch / MLP_TR), dim3(MLP_BS), args, 0, s));
        } else {
            for (uint32_t e = 0; e < ne; e++) {
                hipLaunchKernelGGL(mlp_gather_kernel, dim3(std::min<uint32_t>((ntr * (MLP_IN / 4) + 255) / 256, 2048)),
                                   dim3(256), 0, s, B.Xs, B.Y, B.perm + (size_t)e * ntr, ntr, B.Xp, B.Yp);
                for (uint32_t i = 0; i < spe; i++) {
                    const uint32_t nr = steps_rows[i];
                    rc = mlp_pass(s, B.P, B.Xp, B.Yp, nullptr, i * cfg.batch, nr,
                                  sched[(size_t)e * spe + i].inv_rows, nullptr, B.G, B.lsum, true);
                    if (rc) return rc;
                    hipLaunchKernelGGL(mlp_step_adam_kernel, dim3((MLP_NCOLS + MLP_STEP_COLS - 1) / MLP_STEP_COLS),
                                       dim3(MLP_BS), 0, s, B.P, B.M, B.V, B.G, (nr + MLP_TR - 1) / MLP_TR,
                                       sched[(size_t)e * spe + i], b1, b2, eps);
                }
            }
            TK_HIP(hipGetLastError());
        }
        TK_HIP(hipEventRecord(ev[1], s));
        TK_HIP(hipStreamSynchronize(s));
        float ms = 0.0f;
        TK_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        ms_steps += ms;
        nsteps += ne * spe;
        for (uint32_t i = 0; i < spe; i++) flops += (double)ne * steps_rows[i] * 6.0 * MLP_NP;
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
        rep->steps = nsteps;
        rep->flops = flops;
        rep->ms_steps = ms_steps;
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
    const uint32_t g = (nq + MLP_TR - 1) / MLP_TR;
    hipLaunchKernelGGL(mlp_tau_kernel, dim3(g), dim3(MLP_BS), 0, ctx->stream, L->d_params, B.q, nq, L->sk, L->margin,
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
    const uint32_t g = (nb + MLP_TR - 1) / MLP_TR;
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
    hipStream_t s = ctx->stream;
    int rc;
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
