/*
 * CPU oracle for the top-k query hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference's CPU algorithms for the
 * hot path (avionicscode/TopK, header-only C++ under cpu/, gpu/, common/).
 * It is the checker the HIP path is compared against; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in the product (topk_amd/) links, loads or calls it.
 *
 * Build flags are part of the contract: -O2 -ffp-contract=off, never
 * -ffast-math, so every product x*w and every partial sum is rounded to
 * binary32 separately, in the query's attribute order.
 *
 * Ranking contract (restated from cpu/TPAc.h:118-207 and cpu/AA.h:53-58,
 * with the id-aware tie-break the reference lacks, cpu/TPAc.h:151-185):
 *   score_i = fl(...fl(fl(0 + fl(x[a0][i]*w[a0])) + fl(x[a1][i]*w[a1]))...)
 *   rank objects by (score desc, id asc); the top-k are the first k.
 * Each (score, id) is folded into one 64-bit key that is larger exactly
 * when the object ranks higher:
 *   key = okey(score) << 32 | (0xFFFFFFFF - id)
 * where okey maps binary32 bit patterns to uint32 monotonically.
 *
 * Parity status of each function is listed in its comment; every function
 * is pinned by the CPU tests in tests/test_oracle.py (closed forms, brute
 * force on tiny inputs, invariants).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#if defined(__AVX2__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------ */
/* keys                                                                */
/* ------------------------------------------------------------------ */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Order-preserving binary32 -> uint32 map (negative floats reversed). */
uint32_t or_okey(float s)
{
    uint32_t u = f2u(s);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

float or_okey_inv(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return u2f(u);
}

uint64_t or_key(float s, uint64_t id)
{
    return ((uint64_t)or_okey(s) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)id);
}

void or_key_decode(uint64_t key, uint64_t *id, float *s)
{
    *id = (uint64_t)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu));
    *s = or_okey_inv((uint32_t)(key >> 32));
}

/* ------------------------------------------------------------------ */
/* scoring: cpu/TPAc.h:129-142 (weights indexed by attribute id,       */
/* products rounded then added left to right, accumulator starts at 0) */
/* ------------------------------------------------------------------ */

static inline float score_one(const float *cols, uint64_t ld, uint64_t i,
                              uint32_t qq, const float *w, const uint8_t *attr)
{
    float acc = 0.0f;
    for (uint32_t m = 0; m < qq; m++) {
        float p = cols[(uint64_t)attr[m] * ld + i] * w[attr[m]];
        acc = acc + p;
    }
    return acc;
}

void or_score(const float *cols, uint64_t n, uint64_t ld, uint32_t qq,
              const float *w, const uint8_t *attr, float *out)
{
    for (uint64_t i = 0; i < n; i++) out[i] = score_one(cols, ld, i, qq, w, attr);
}

/* Real-data scoring (the reference's LD == 2 build, cpu/TPAc.h:143-146,
 * 237-240, 291-294): after every term the running score is divided by qq,
 * each step rounded to binary32.  Parity pinned by exact powers of two
 * (qq = 2, 4: every division is exact) and a float32 step-by-step
 * restatement for other qq (tests/test_oracle.py). */
static inline float score_one_div(const float *cols, uint64_t ld, uint64_t i,
                                  uint32_t qq, const float *w, const uint8_t *attr)
{
    const float q = (float)qq;
    float acc = 0.0f;
    for (uint32_t m = 0; m < qq; m++) {
        float p = cols[(uint64_t)attr[m] * ld + i] * w[attr[m]];
        acc = acc + p;
        acc = acc / q;
    }
    return acc;
}

void or_score_div(const float *cols, uint64_t n, uint64_t ld, uint32_t qq,
                  const float *w, const uint8_t *attr, float *out)
{
    for (uint64_t i = 0; i < n; i++) out[i] = score_one_div(cols, ld, i, qq, w, attr);
}

/* ------------------------------------------------------------------ */
/* size-k min-heap on 64-bit keys (cpu/AA.h:53-58 MaxCMP semantics:   */
/* the top is the worst kept element; replace only if strictly better) */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t *a; uint32_t n, cap; } kheap;

static void heap_sift_down(uint64_t *a, uint32_t n, uint32_t i)
{
    for (;;) {
        uint32_t l = 2 * i + 1, r = l + 1, s = i;
        if (l < n && a[l] < a[s]) s = l;
        if (r < n && a[r] < a[s]) s = r;
        if (s == i) return;
        uint64_t t = a[i]; a[i] = a[s]; a[s] = t; i = s;
    }
}

static void heap_push(kheap *h, uint64_t key)
{
    if (h->n < h->cap) {
        uint32_t i = h->n++;
        h->a[i] = key;
        while (i > 0) {
            uint32_t p = (i - 1) / 2;
            if (h->a[p] <= h->a[i]) break;
            uint64_t t = h->a[p]; h->a[p] = h->a[i]; h->a[i] = t; i = p;
        }
    } else if (key > h->a[0]) {
        h->a[0] = key;
        heap_sift_down(h->a, h->n, 0);
    }
}

static int cmp_u64_desc(const void *x, const void *y)
{
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return (a < b) - (a > b);
}

/* Sort the kept keys descending and decode them.  Returns the count. */
static uint32_t heap_emit(kheap *h, uint64_t *ids, float *scores, float *thr)
{
    qsort(h->a, h->n, sizeof(uint64_t), cmp_u64_desc);
    for (uint32_t j = 0; j < h->n; j++) or_key_decode(h->a[j], &ids[j], &scores[j]);
    if (thr) *thr = h->n ? scores[h->n - 1] : -INFINITY;
    return h->n;
}

/*
 * Scalar FTE top-k: cpu/TPAc.h:35-115 / 118-207 restated (one pass, one
 * heap), with the id-aware key.  Outputs (score desc, id asc); returns
 * min(k, n).  Parity pinned by brute-force sort and closed forms.
 */
uint32_t or_fte_topk(const float *cols, uint64_t n, uint64_t ld, uint32_t k,
                     uint32_t qq, const float *w, const uint8_t *attr,
                     uint64_t *ids, float *scores, float *thr)
{
    kheap h = { (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1)), 0, k };
    for (uint64_t i = 0; i < n; i++)
        heap_push(&h, or_key(score_one(cols, ld, i, qq, w, attr), i));
    uint32_t c = heap_emit(&h, ids, scores, thr);
    free(h.a);
    return c;
}

/*
 * 64-bit ids (stores of 2^32 - 1 objects or more; the reference's tuples
 * carry uint64_t tids, cpu/AA.h:29-35): the same ranking with the id in 64
 * bits, key = okey(score) << 64 | (2^64 - 1 - id) in a 128-bit integer.
 * Object i of the columns has id id_base + i.  Threads own contiguous
 * ranges with private heaps, merged at the end, as or_tpac_threads.
 * Parity pinned by agreement with or_tpac_threads below 2^32 and by a
 * brute-force lexicographic sort of ranges that cross 2^32.
 */
typedef unsigned __int128 u128;
typedef struct { u128 *a; uint32_t n, cap; } wheap;

static void wheap_push(wheap *h, u128 key)
{
    if (h->n < h->cap) {
        uint32_t i = h->n++;
        h->a[i] = key;
        while (i > 0) {
            uint32_t p = (i - 1) / 2;
            if (h->a[p] <= h->a[i]) break;
            u128 t = h->a[p]; h->a[p] = h->a[i]; h->a[i] = t; i = p;
        }
    } else if (key > h->a[0]) {
        h->a[0] = key;
        for (uint32_t i = 0;;) {
            uint32_t l = 2 * i + 1, r = l + 1, m = i;
            if (l < h->n && h->a[l] < h->a[m]) m = l;
            if (r < h->n && h->a[r] < h->a[m]) m = r;
            if (m == i) break;
            u128 t = h->a[i]; h->a[i] = h->a[m]; h->a[m] = t; i = m;
        }
    }
}

static inline u128 wkey(float s, uint64_t id)
{
    return ((u128)or_okey(s) << 64) | (u128)(UINT64_MAX - id);
}

static int cmp_u128_desc(const void *x, const void *y)
{
    u128 a = *(const u128 *)x, b = *(const u128 *)y;
    return (a < b) - (a > b);
}

uint32_t or_fte_topk_wide(const float *cols, uint64_t n, uint64_t ld, uint32_t k,
                          uint32_t qq, const float *w, const uint8_t *attr, uint64_t id_base,
                          int threads, uint64_t *ids, float *scores, float *thr)
{
    if (threads < 1) threads = 1;
    u128 *all = (u128 *)malloc(sizeof(u128) * (uint64_t)(k ? k : 1) * threads);
    uint32_t *cnt = (uint32_t *)calloc(threads, sizeof(uint32_t));
#pragma omp parallel num_threads(threads)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        wheap h = { all + (uint64_t)k * t, 0, k };
        uint64_t lo = (uint64_t)t * (n / threads) + ((uint64_t)t < n % threads ? (uint64_t)t : n % threads);
        uint64_t hi = lo + n / threads + ((uint64_t)t < n % threads ? 1 : 0);
        for (uint64_t i = lo; i < hi; i++)
            wheap_push(&h, wkey(score_one(cols, ld, i, qq, w, attr), id_base + i));
        cnt[t] = h.n;
    }
    wheap g = { (u128 *)malloc(sizeof(u128) * (k ? k : 1)), 0, k };
    for (int t = 0; t < threads; t++)
        for (uint32_t j = 0; j < cnt[t]; j++) wheap_push(&g, all[(uint64_t)k * t + j]);
    qsort(g.a, g.n, sizeof(u128), cmp_u128_desc);
    for (uint32_t j = 0; j < g.n; j++) {
        ids[j] = UINT64_MAX - (uint64_t)g.a[j];
        scores[j] = or_okey_inv((uint32_t)(g.a[j] >> 64));
    }
    if (thr) *thr = g.n ? scores[g.n - 1] : -INFINITY;
    uint32_t c = g.n;
    free(g.a); free(all); free(cnt);
    return c;
}

/* or_fte_topk over real-data scores (score_one_div). */
uint32_t or_fte_topk_div(const float *cols, uint64_t n, uint64_t ld, uint32_t k,
                         uint32_t qq, const float *w, const uint8_t *attr,
                         uint64_t *ids, float *scores, float *thr)
{
    kheap h = { (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1)), 0, k };
    for (uint64_t i = 0; i < n; i++)
        heap_push(&h, or_key(score_one_div(cols, ld, i, qq, w, attr), i));
    uint32_t c = heap_emit(&h, ids, scores, thr);
    free(h.a);
    return c;
}

/*
 * SIMD + threads FTE baseline: cpu/TPAc.h:256-358 restated.  Each thread
 * owns a contiguous object range, scores 16 objects per step with two
 * 8-wide mul/add (separately rounded, so bit-identical to or_fte_topk),
 * keeps a private heap, and the heaps are merged at the end
 * (cpu/TPAc.h:335-346).  Used as the timed CPU baseline.
 */
uint32_t or_tpac_threads(const float *cols, uint64_t n, uint64_t ld, uint32_t k,
                         uint32_t qq, const float *w, const uint8_t *attr,
                         int threads, uint64_t *ids, float *scores, float *thr)
{
    if (threads < 1) threads = 1;
    uint64_t *all = (uint64_t *)malloc(sizeof(uint64_t) * (uint64_t)k * threads);
    uint32_t *cnt = (uint32_t *)calloc(threads, sizeof(uint32_t));
#pragma omp parallel num_threads(threads)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        kheap h = { all + (uint64_t)k * t, 0, k };
        uint64_t lo = (uint64_t)t * n / threads, hi = (uint64_t)(t + 1) * n / threads;
        uint64_t i = lo;
#if defined(__AVX2__)
        float sc[16] __attribute__((aligned(32)));
        for (; i + 16 <= hi; i += 16) {
            __m256 s0 = _mm256_setzero_ps(), s1 = _mm256_setzero_ps();
            for (uint32_t m = 0; m < qq; m++) {
                const float *c = cols + (uint64_t)attr[m] * ld + i;
                __m256 wv = _mm256_set1_ps(w[attr[m]]);
                s0 = _mm256_add_ps(s0, _mm256_mul_ps(_mm256_loadu_ps(c), wv));
                s1 = _mm256_add_ps(s1, _mm256_mul_ps(_mm256_loadu_ps(c + 8), wv));
            }
            _mm256_store_ps(sc, s0);
            _mm256_store_ps(sc + 8, s1);
            for (int l = 0; l < 16; l++) heap_push(&h, or_key(sc[l], i + l));
        }
#endif
        for (; i < hi; i++) heap_push(&h, or_key(score_one(cols, ld, i, qq, w, attr), i));
        cnt[t] = h.n;
    }
    kheap g = { (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1)), 0, k };
    for (int t = 0; t < threads; t++)
        for (uint32_t j = 0; j < cnt[t]; j++) heap_push(&g, all[(uint64_t)k * t + j]);
    uint32_t c = heap_emit(&g, ids, scores, thr);
    free(g.a); free(all); free(cnt);
    return c;
}

/*
 * Multi-query baseline: cpu/TPAc.h:209-254 + validation/bench_cpu.h:198-215
 * restated: one full scan per query, queries spread over threads.  Query q
 * uses attr[q*attr_stride .. +qq[q]), weights w[q*w_stride ..], k[q]; its
 * results go to ids/scores at q*k_stride.  Returns 0.
 */
int or_mq_threads(const float *cols, uint64_t n, uint64_t ld, uint32_t nq,
                  const uint32_t *k, const uint32_t *qq, const float *w, uint32_t w_stride,
                  const uint8_t *attr, uint32_t attr_stride, int threads,
                  uint64_t *ids, float *scores, float *thr, uint32_t k_stride)
{
    if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
        or_fte_topk(cols, n, ld, k[q], qq[q], w + (uint64_t)q * w_stride,
                    attr + (uint64_t)q * attr_stride, ids + (uint64_t)q * k_stride,
                    scores + (uint64_t)q * k_stride, thr + q);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* radix select: gpu/radix_select.h:73-120 and gpu/GPA.h:116-158        */
/* restated.  MSD over 32-bit patterns, 4-bit digits, masks            */
/* 0xF0000000 >> 4i; returns the element of ascending rank r (0-based) */
/* ------------------------------------------------------------------ */

uint32_t or_radix_select_u32(const uint32_t *keys, uint64_t n, uint64_t r)
{
    uint32_t prefix = 0, prefix_mask = 0, digit_mask = 0xF0000000u;
    uint32_t shf = 28;
    uint64_t rem = r;
    for (int pass = 0; pass < 8; pass++) {
        uint64_t bins[16] = { 0 };
        for (uint64_t i = 0; i < n; i++)
            if ((keys[i] & prefix_mask) == prefix) bins[(keys[i] & digit_mask) >> shf]++;
        uint32_t digit = 15;
        uint64_t cum = 0;
        for (uint32_t b = 0; b < 16; b++) {
            if (cum + bins[b] > rem) { digit = b; rem -= cum; break; }
            cum += bins[b];
        }
        prefix |= digit << shf;
        prefix_mask |= digit_mask;
        digit_mask >>= 4;
        shf -= 4;
    }
    return prefix;
}

/* Same selection over 64-bit keys (16 passes of 4 bits). */
uint64_t or_radix_select_u64(const uint64_t *keys, uint64_t n, uint64_t r)
{
    uint64_t prefix = 0, prefix_mask = 0, digit_mask = 0xF000000000000000ull;
    uint32_t shf = 60;
    uint64_t rem = r;
    for (int pass = 0; pass < 16; pass++) {
        uint64_t bins[16] = { 0 };
        for (uint64_t i = 0; i < n; i++)
            if ((keys[i] & prefix_mask) == prefix) bins[(keys[i] & digit_mask) >> shf]++;
        uint64_t digit = 15, cum = 0;
        for (uint32_t b = 0; b < 16; b++) {
            if (cum + bins[b] > rem) { digit = b; rem -= cum; break; }
            cum += bins[b];
        }
        prefix |= digit << shf;
        prefix_mask |= digit_mask;
        digit_mask >>= 4;
        shf -= 4;
    }
    return prefix;
}

/* ------------------------------------------------------------------ */
/* normalisation: common/common.h:29-60 restated (per column, max      */
/* starts at 0 and min at FLT_MAX; x = (x - min) / (max - min))        */
/* ------------------------------------------------------------------ */

void or_normalize(float *cols, uint64_t n, uint32_t d, uint64_t ld)
{
    for (uint32_t m = 0; m < d; m++) {
        float mx = 0.0f, mn = 3.40282347e+38f;
        float *c = cols + (uint64_t)m * ld;
        for (uint64_t i = 0; i < n; i++) {
            float v = c[i];
            mx = (mx < v) ? v : mx;       /* std::max(mx, v) */
            mn = (v < mn) ? v : mn;       /* std::min(mn, v) */
        }
        float mm = mx - mn;
        for (uint64_t i = 0; i < n; i++) c[i] = (c[i] - mn) / mm;
    }
}

/* ------------------------------------------------------------------ */
/* reorder_attr: gpu/reorder_attr.h:13-236 / cpu/reorder_attr_cpu_c.h   */
/* restated.  Per object, its d values are sorted descending across the */
/* columns, in place (column 0 holds each object's maximum).           */
/* Invariant: gpu/GPA.h:96-114 check_order.                             */
/* ------------------------------------------------------------------ */

void or_reorder_attr(float *cols, uint64_t n, uint32_t d, uint64_t ld)
{
    float v[64];
    for (uint64_t i = 0; i < n; i++) {
        for (uint32_t m = 0; m < d; m++) v[m] = cols[(uint64_t)m * ld + i];
        for (uint32_t a = 1; a < d; a++) {          /* insertion sort, descending */
            float x = v[a];
            int b = (int)a - 1;
            while (b >= 0 && v[b] < x) { v[b + 1] = v[b]; b--; }
            v[b + 1] = x;
        }
        for (uint32_t m = 0; m < d; m++) cols[(uint64_t)m * ld + i] = v[m];
    }
}

/* Returns the first object index that violates descending order, or n. */
uint64_t or_check_order(const float *cols, uint64_t n, uint32_t d, uint64_t ld)
{
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t m = 0; m + 1 < d; m++)
            if (!(cols[(uint64_t)m * ld + i] >= cols[(uint64_t)(m + 1) * ld + i])) return i;
    return n;
}

/* ------------------------------------------------------------------ */
/* BTA: gpu/BTA.h:36-175 restated at block granularity.  For each block */
/* of `block` objects, the block's top-k keys in descending order (the  */
/* reference carries scores only; here each score travels with its id). */
/* out has nblocks*k entries; blocks with fewer than k objects are      */
/* padded with key 0.                                                   */
/* ------------------------------------------------------------------ */

void or_bta_blocks(const float *cols, uint64_t n, uint64_t ld, uint32_t qq,
                   const float *w, const uint8_t *attr, uint32_t block, uint32_t k,
                   uint64_t *out)
{
    uint64_t nb = (n + block - 1) / block;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * block);
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t lo = b * block, hi = lo + block < n ? lo + block : n;
        uint32_t c = 0;
        for (uint64_t i = lo; i < hi; i++) tmp[c++] = or_key(score_one(cols, ld, i, qq, w, attr), i);
        qsort(tmp, c, sizeof(uint64_t), cmp_u64_desc);
        for (uint32_t j = 0; j < k; j++) out[b * k + j] = j < c ? tmp[j] : 0;
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* GPA prune + compact: gpu/prune_tuples.h:10-29,90-118 and the driver  */
/* loop gpu/GPA.h:207-220 restated for one iteration.  Keeps candidate  */
/* j if fl(score[j] + fl(colC[id[j]] * suffix_len)) >= thr, compacts    */
/* (stable) and adds the next column: score += colN[id].  Returns the   */
/* survivor count.                                                      */
/* ------------------------------------------------------------------ */

uint64_t or_gpa_step(uint32_t *ids, float *score, uint64_t cnt, const float *colC,
                     const float *colN, float thr, float suffix_len)
{
    uint64_t o = 0;
    for (uint64_t j = 0; j < cnt; j++) {
        uint32_t id = ids[j];
        float ub = score[j] + colC[id] * suffix_len;
        if (ub >= thr) {
            ids[o] = id;
            score[o] = colN ? score[j] + colN[id] : score[j];
            o++;
        }
    }
    return o;
}

/* ------------------------------------------------------------------ */
/* PTA: cpu/PTA.h:357-422 / 569-671 restated over an explicit partition */
/* layout.  The layout is an input, so CPU and GPU scan identical        */
/* partitions: partition p occupies positions [poff[p], poff[p]+psz[p])  */
/* of the partition-order columns (pdata[m*ld + j] is attribute m of the */
/* object at position j, pid[j] its global id).  A partition is skipped  */
/* when its bound cannot beat the k-th kept key (strict: skip iff        */
/* okey(bound) < okey(kth score) with a full heap); the reference stops  */
/* on >= (cpu/PTA.h:340), which is unsafe once ties are broken by id.    */
/* The bound is the query's score evaluated at the partition's           */
/* per-attribute maximum (minimum where the weight is negative), which   */
/* is FP-monotone.  Partitions are visited in descending bound order.    */
/* Returns the count and writes the number of scored partitions.         */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t key; uint32_t p; } pbound;

static int cmp_pbound_desc(const void *x, const void *y)
{
    const pbound *a = (const pbound *)x, *b = (const pbound *)y;
    if (a->key != b->key) return (a->key < b->key) ? 1 : -1;
    return (a->p > b->p) - (a->p < b->p);
}

float or_pta_bound(const float *pmax, const float *pmin, uint32_t d, uint64_t p,
                   uint32_t qq, const float *w, const uint8_t *attr)
{
    float acc = 0.0f;
    for (uint32_t m = 0; m < qq; m++) {
        uint32_t a = attr[m];
        float x = (w[a] >= 0.0f) ? pmax[p * d + a] : pmin[p * d + a];
        float pr = x * w[a];
        acc = acc + pr;
    }
    return acc;
}

static pbound *pta_sorted_bounds(uint32_t nparts, const uint32_t *psz, const float *pmax,
                                 const float *pmin, uint32_t d, uint32_t qq, const float *w,
                                 const uint8_t *attr, uint32_t *nonempty)
{
    pbound *pb = (pbound *)malloc(sizeof(pbound) * (nparts ? nparts : 1));
    uint32_t c = 0;
    for (uint32_t p = 0; p < nparts; p++) {
        if (!psz[p]) continue;
        pb[c].key = (uint64_t)or_okey(or_pta_bound(pmax, pmin, d, p, qq, w, attr)) << 32;
        pb[c].p = p;
        c++;
    }
    qsort(pb, c, sizeof(pbound), cmp_pbound_desc);
    *nonempty = c;
    return pb;
}

uint32_t or_pta_topk(const float *pdata, const uint32_t *pid, uint64_t ld, uint32_t d,
                     uint32_t nparts, const uint32_t *poff, const uint32_t *psz,
                     const float *pmax, const float *pmin,
                     uint32_t k, uint32_t qq, const float *w, const uint8_t *attr,
                     uint64_t *ids, float *scores, float *thr, uint64_t *scored)
{
    uint32_t np;
    pbound *pb = pta_sorted_bounds(nparts, psz, pmax, pmin, d, qq, w, attr, &np);
    kheap h = { (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1)), 0, k };
    uint64_t sc = 0;
    for (uint32_t t = 0; t < np; t++) {
        uint32_t bkey = (uint32_t)(pb[t].key >> 32);
        if (h.n == k && bkey < (uint32_t)(h.a[0] >> 32)) break;   /* strict skip */
        uint32_t p = pb[t].p;
        for (uint64_t j = poff[p]; j < (uint64_t)poff[p] + psz[p]; j++)
            heap_push(&h, or_key(score_one(pdata, ld, j, qq, w, attr), pid[j]));
        sc++;
    }
    if (scored) *scored = sc;
    uint32_t c = heap_emit(&h, ids, scores, thr);
    free(h.a); free(pb);
    return c;
}

/* Threaded PTA baseline: partitions are dealt to threads in descending  */
/* bound order (cpu/PTA.h:569-671 marches partitions per thread); each   */
/* thread prunes against its own heap, then the heaps merge.             */
uint32_t or_pta_threads(const float *pdata, const uint32_t *pid, uint64_t ld, uint32_t d,
                        uint32_t nparts, const uint32_t *poff, const uint32_t *psz,
                        const float *pmax, const float *pmin,
                        uint32_t k, uint32_t qq, const float *w, const uint8_t *attr,
                        int threads, uint64_t *ids, float *scores, float *thr, uint64_t *scored)
{
    if (threads < 1) threads = 1;
    uint32_t np;
    pbound *pb = pta_sorted_bounds(nparts, psz, pmax, pmin, d, qq, w, attr, &np);
    uint64_t *all = (uint64_t *)malloc(sizeof(uint64_t) * (uint64_t)k * threads);
    uint32_t *cnt = (uint32_t *)calloc(threads, sizeof(uint32_t));
    uint64_t *scnt = (uint64_t *)calloc(threads, sizeof(uint64_t));
#pragma omp parallel num_threads(threads)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        kheap h = { all + (uint64_t)k * t, 0, k };
        for (uint32_t r = t; r < np; r += threads) {
            uint32_t bkey = (uint32_t)(pb[r].key >> 32);
            if (h.n == k && bkey < (uint32_t)(h.a[0] >> 32)) break;
            uint32_t p = pb[r].p;
            for (uint64_t j = poff[p]; j < (uint64_t)poff[p] + psz[p]; j++)
                heap_push(&h, or_key(score_one(pdata, ld, j, qq, w, attr), pid[j]));
            scnt[t]++;
        }
        cnt[t] = h.n;
    }
    kheap g = { (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1)), 0, k };
    uint64_t sc = 0;
    for (int t = 0; t < threads; t++) {
        sc += scnt[t];
        for (uint32_t j = 0; j < cnt[t]; j++) heap_push(&g, all[(uint64_t)k * t + j]);
    }
    if (scored) *scored = sc;
    uint32_t c = heap_emit(&g, ids, scores, thr);
    free(g.a); free(all); free(cnt); free(scnt); free(pb);
    return c;
}

/* ------------------------------------------------------------------ */
/* MLP 96 -> 64 -> 64 -> 1 (ReLU), no reference counterpart.  Float64   */
/* forward and backward of the mean squared error over a batch, used to */
/* check the HIP training kernels' gradients.  Parameter layout:        */
/* W1[64][96], b1[64], W2[64][64], b2[64], W3[64], b3[1] (10433 values).*/
/* ------------------------------------------------------------------ */

#define MLP_IN 96
#define MLP_H 64

static void mlp_offsets(uint32_t *o)
{
    o[0] = 0;                        /* W1 */
    o[1] = o[0] + MLP_H * MLP_IN;    /* b1 */
    o[2] = o[1] + MLP_H;             /* W2 */
    o[3] = o[2] + MLP_H * MLP_H;     /* b2 */
    o[4] = o[3] + MLP_H;             /* W3 */
    o[5] = o[4] + MLP_H;             /* b3 */
    o[6] = o[5] + 1;                 /* total */
}

uint32_t or_mlp_nparams(void) { uint32_t o[7]; mlp_offsets(o); return o[6]; }

/* out[b] = prediction for row b of x (batch x 96). */
void or_mlp_forward(const double *p, const double *x, uint32_t batch, double *out)
{
    uint32_t o[7]; mlp_offsets(o);
    double h1[MLP_H], h2[MLP_H];
    for (uint32_t b = 0; b < batch; b++) {
        const double *xb = x + (uint64_t)b * MLP_IN;
        for (int j = 0; j < MLP_H; j++) {
            double s = p[o[1] + j];
            for (int i = 0; i < MLP_IN; i++) s += p[o[0] + j * MLP_IN + i] * xb[i];
            h1[j] = s > 0 ? s : 0;
        }
        for (int j = 0; j < MLP_H; j++) {
            double s = p[o[3] + j];
            for (int i = 0; i < MLP_H; i++) s += p[o[2] + j * MLP_H + i] * h1[i];
            h2[j] = s > 0 ? s : 0;
        }
        double s = p[o[5]];
        for (int i = 0; i < MLP_H; i++) s += p[o[4] + i] * h2[i];
        out[b] = s;
    }
}

/* Loss = mean_b (pred_b - y_b)^2; grad has nparams entries.  Returns loss. */
double or_mlp_backward(const double *p, const double *x, const double *y, uint32_t batch,
                       double *grad)
{
    uint32_t o[7]; mlp_offsets(o);
    memset(grad, 0, sizeof(double) * o[6]);
    double a1[MLP_H], h1[MLP_H], a2[MLP_H], h2[MLP_H], g2[MLP_H], g1[MLP_H];
    double loss = 0;
    for (uint32_t b = 0; b < batch; b++) {
        const double *xb = x + (uint64_t)b * MLP_IN;
        for (int j = 0; j < MLP_H; j++) {
            double s = p[o[1] + j];
            for (int i = 0; i < MLP_IN; i++) s += p[o[0] + j * MLP_IN + i] * xb[i];
            a1[j] = s; h1[j] = s > 0 ? s : 0;
        }
        for (int j = 0; j < MLP_H; j++) {
            double s = p[o[3] + j];
            for (int i = 0; i < MLP_H; i++) s += p[o[2] + j * MLP_H + i] * h1[i];
            a2[j] = s; h2[j] = s > 0 ? s : 0;
        }
        double s = p[o[5]];
        for (int i = 0; i < MLP_H; i++) s += p[o[4] + i] * h2[i];
        double e = s - y[b];
        loss += e * e;
        double go = 2.0 * e / batch;
        grad[o[5]] += go;
        for (int i = 0; i < MLP_H; i++) {
            grad[o[4] + i] += go * h2[i];
            g2[i] = (a2[i] > 0) ? go * p[o[4] + i] : 0;
        }
        for (int j = 0; j < MLP_H; j++) {
            grad[o[3] + j] += g2[j];
            for (int i = 0; i < MLP_H; i++) grad[o[2] + j * MLP_H + i] += g2[j] * h1[i];
        }
        for (int i = 0; i < MLP_H; i++) {
            double s2 = 0;
            for (int j = 0; j < MLP_H; j++) s2 += g2[j] * p[o[2] + j * MLP_H + i];
            g1[i] = (a1[i] > 0) ? s2 : 0;
        }
        for (int j = 0; j < MLP_H; j++) {
            grad[o[1] + j] += g1[j];
            for (int i = 0; i < MLP_IN; i++) grad[o[0] + j * MLP_IN + i] += g1[j] * xb[i];
        }
    }
    return loss / batch;
}

/* ------------------------------------------------------------------ */
/* text loader (File<T>, input/File.h)                                 */
/* ------------------------------------------------------------------ */
#include <stdio.h>

/* File<T>::count (input/File.h:107-128): d = commas on the first line + 1,
   n = number of '\n' bytes in the file. */
int or_csv_count(const char *path, uint64_t *n, uint32_t *d)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    uint64_t nl = 0;
    uint32_t dd = 1;
    int first = 1, c;
    while ((c = fgetc(f)) != EOF) {
        if (first && c == ',') dd++;
        if (c == '\n') { nl++; first = 0; }
    }
    fclose(f);
    *n = nl;
    *d = dd;
    return 0;
}

/* File<T>::read_scanf_t / read_scanf (input/File.h:233-295): rows read
   with fscanf and the "%f,%f,...,%f" specifier of line_specifier
   (:86-104), stored x[m*n+i] (colmajor) or x[i*d+m], until n rows are read
   or a row's first value fails.  Returns the rows read, or -1 for a row
   that ended after fewer than d values (the reference keeps the previous
   row's values there) and -2 if the file cannot be opened. */
int64_t or_csv_read(const char *path, uint64_t n, uint32_t d, int colmajor, float *out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    uint64_t i = 0;
    float v[64];
    while (i < n) {
        int got = fscanf(f, "%f", &v[0]);
        if (got != 1) break;
        uint32_t m = 1;
        for (; m < d; m++)
            if (fscanf(f, ",%f", &v[m]) != 1) break;
        if (m < d) { fclose(f); return -1; }
        for (m = 0; m < d; m++) {
            if (colmajor) out[(uint64_t)m * n + i] = v[m];
            else out[i * d + m] = v[m];
        }
        i++;
    }
    fclose(f);
    return (int64_t)i;
}
