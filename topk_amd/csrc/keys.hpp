// Ranking keys and the synthetic data generator, shared by host and device
// code.  Both are pure integer / IEEE binary32 arithmetic compiled with
// -ffp-contract=off, so host and gfx950 produce identical bits.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define TK_HD __host__ __device__ __forceinline__
#else
#define TK_HD inline
#endif

namespace topk {

TK_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
TK_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// Monotone binary32 -> uint32 map: a < b (as floats, no NaN) <=> okey(a) < okey(b).
// +0 and -0 map to different keys; scores start from +0 and an IEEE sum can
// only be -0 when both operands are -0, so a score is never -0.
TK_HD uint32_t okey(float s) {
    uint32_t u = f2u(s);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
TK_HD float okey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return u2f(u);
}
// 64-bit ranking key: larger key <=> ranks higher under (score desc, id asc).
// Key 0 is below every real object and marks empty slots.
TK_HD uint64_t rkey(float s, uint32_t id) {
    return ((uint64_t)okey(s) << 32) | (uint64_t)(0xFFFFFFFFu - id);
}
TK_HD uint64_t rkey_from_okey(uint32_t ok, uint32_t id) {
    return ((uint64_t)ok << 32) | (uint64_t)(0xFFFFFFFFu - id);
}
TK_HD uint32_t rkey_id(uint64_t k) { return 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu); }
TK_HD float rkey_score(uint64_t k) { return okey_inv((uint32_t)(k >> 32)); }

// ---------------------------------------------------------------------------
// Synthetic generator (the benchmark distributions).  Counter-based: the value of
// (object i, attribute m) depends only on (seed, i, m), so any object range
// can be generated independently, on the host or on the device.
// ---------------------------------------------------------------------------

TK_HD uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser (Steele et al., 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 24 random bits for (object, slot, tag); slot 0..62 = attribute, 63 = per-object.
TK_HD uint32_t rbits24(uint64_t sk, uint64_t i, uint32_t slot, uint32_t tag) {
    return (uint32_t)(mix64(sk + ((i * 64u + slot) * 16u + tag)) >> 40);
}
// U[0,1) with 24-bit resolution (exact in binary32).
TK_HD float runif(uint64_t sk, uint64_t i, uint32_t slot, uint32_t tag) {
    return (float)rbits24(sk, i, slot, tag) * 5.9604644775390625e-08f;  // 2^-24
}
// Approximate N(0,1) by Irwin-Hall: sum of 12 uniforms minus 6, summed in
// integers, then one rounding to binary32.
TK_HD float rnorm(uint64_t sk, uint64_t i, uint32_t slot) {
    int64_t s = 0;
    for (uint32_t t = 1; t <= 12; t++) s += (int64_t)rbits24(sk, i, slot, t);
    s -= 6ll * 16777216ll;
    return (float)s * 5.9604644775390625e-08f;
}
TK_HD float clip01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

enum { GEN_CORRELATED = 0, GEN_INDEPENDENT = 1, GEN_ANTICORRELATED = 2, GEN_MIXED = 3 };

// Writes the d attributes of object i (of an n-object dataset) into v[0..d).
TK_HD void gen_object(float *v, uint64_t n, uint32_t d, int distr, uint64_t seed, uint64_t i) {
    const uint64_t sk = mix64(seed ^ 0x5DEECE66Dull);
    int dist = distr;
    if (distr == GEN_MIXED) dist = (i < n / 3) ? GEN_CORRELATED : (i < 2 * (n / 3) ? GEN_INDEPENDENT : GEN_ANTICORRELATED);
    if (dist == GEN_INDEPENDENT) {
        for (uint32_t m = 0; m < d; m++) v[m] = runif(sk, i, m, 0);
    } else if (dist == GEN_CORRELATED) {
        float u = runif(sk, i, 63, 0);
        for (uint32_t m = 0; m < d; m++) {
            float z = rnorm(sk, i, m) * 0.05f;
            v[m] = clip01(u + z);
        }
    } else {  // anticorrelated: shift a uniform point onto sum = d/2 + N(0,0.05), clip
        float s = 0.0f;
        for (uint32_t m = 0; m < d; m++) { v[m] = runif(sk, i, m, 0); s = s + v[m]; }
        float z = rnorm(sk, i, 63) * 0.05f;
        float target = (float)d * 0.5f + z;
        float c = (target - s) / (float)d;
        for (uint32_t m = 0; m < d; m++) v[m] = clip01(v[m] + c);
    }
}

}  // namespace topk
