// Upper bound of a query's score over one partition, shared by the PTA and
// batch paths.
//
// A partition stores per-attribute max/min, the largest attribute sum of its
// objects (rounded up) and the sum of its per-attribute minima (rounded
// down).  The bound is the maximum of sum_a W_a x_a over the partition's box
// under the attribute-sum limit -- a fractional knapsack: every queried
// attribute starts at its minimum, then the positive weights, heaviest
// first, take the room above their minimum until the sum budget is used --
// capped by the plain box bound, in FP64.  W_a aggregates a weight listed
// twice.  A binary32 score of m terms differs from the exact sum by at most
// about 2m * 2^-24 * sum |w x|; the margin 2^-16 * sum |W| max(|max|,|min|)
// covers that for m <= 32, and the result is rounded up to binary32, so the
// bound is never below an object's computed score.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/topk.h"

namespace topk {

__device__ __forceinline__ float f32_round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
__device__ __forceinline__ float f32_round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

// The distinct queried attributes sorted by aggregated weight, heaviest first.
struct BoundArgs {
    uint32_t nd;
    uint32_t d;
    uint8_t attr[TOPK_MAX_DIMS];
    double w[TOPK_MAX_DIMS];
};

inline BoundArgs make_bound_args(uint32_t d, uint32_t qq, const float *weights, const uint8_t *attr) {
    BoundArgs b{};
    b.d = d;
    double agg[TOPK_MAX_DIMS] = {0};
    bool used[TOPK_MAX_DIMS] = {false};
    for (uint32_t j = 0; j < qq; j++) {
        agg[attr[j]] += (double)weights[attr[j]];
        used[attr[j]] = true;
    }
    uint32_t order[TOPK_MAX_DIMS], nd = 0;
    for (uint32_t a = 0; a < d; a++)
        if (used[a]) order[nd++] = a;
    std::stable_sort(order, order + nd, [&](uint32_t x, uint32_t y) { return agg[x] > agg[y]; });
    b.nd = nd;
    for (uint32_t i = 0; i < nd; i++) {
        b.attr[i] = (uint8_t)order[i];
        b.w[i] = agg[order[i]];
    }
    return b;
}

// The LP bound from operands already in registers: the first nd entries of
// wv (weights, heaviest first), mxv / mnv (their attributes' maxima /
// minima); entries past nd must hold w = 0 and max = min = 0, which add
// exact zeros; the knapsack loop skips them.  psum bounds every object's
// attribute sum from above, pminsum is the sum of all attributes' minima
// rounded down.  The loops are predicated rather than cut short.
template <int NMAX>
__device__ __forceinline__ float lp_bound_regs(const double (&wv)[NMAX], const float (&mxv)[NMAX],
                                               const float (&mnv)[NMAX], float psum, float pminsum) {
    // one pass: every attribute adds w * min to the knapsack value, and a
    // positive weight then takes its room above the minimum while the sum
    // budget lasts (the weights are sorted heaviest first, so the positive
    // ones form a prefix and the budget only shrinks); an attribute's
    // operands are dead once it is added
    double base = 0, box = 0, absw = 0, budget = (double)psum - (double)pminsum;
#pragma unroll
    for (int i = 0; i < NMAX; i++) {
        const double mx = mxv[i], mn = mnv[i], w = wv[i];
        base += w * mn;
        box += w * (w >= 0 ? mx : mn);
        absw += fabs(w) * fmax(fabs(mx), fabs(mn));
        if (w > 0 && budget > 0) {
            const double room = fmin(mx - mn, budget);
            base += w * room;
            budget -= room;
        }
    }
    const double lp = base < box ? base : box;
    return f32_round_up(lp + absw * 1.52587890625e-05 + 1e-30);  // margin 2^-16 * sum |W x|
}

// pmax / pmin point at the partition's per-attribute maxima / minima,
// attribute a at pmax[a * stride] (stride 1: a row; the number of nodes or
// chunks: an attribute-major array, where the lanes of a wave bounding
// consecutive nodes read consecutive words).  NMAX bounds b.nd at compile
// time.  Every operand (weight, maximum, minimum) is read with independent
// loads into registers before any arithmetic, so a bound costs one round
// of memory latency.
template <int NMAX = TOPK_MAX_DIMS>
__device__ __forceinline__ float lp_bound(const BoundArgs &b, const float *__restrict__ pmax,
                                          const float *__restrict__ pmin, float psum, float pminsum,
                                          uint64_t stride = 1) {
    const uint32_t nd = b.nd;
    float mxv[NMAX], mnv[NMAX];
    double wv[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; i++) {
        const bool on = (uint32_t)i < nd;
        const uint32_t ai = on ? b.attr[i] : 0u;  // entries past nd may be stale
        wv[i] = on ? b.w[i] : 0.0;
        mxv[i] = on ? pmax[ai * stride] : 0.0f;
        mnv[i] = on ? pmin[ai * stride] : 0.0f;
    }
    return lp_bound_regs<NMAX>(wv, mxv, mnv, psum, pminsum);
}

}  // namespace topk
