/*
 * topk_amd.hpp -- C++ host mirror of the reference's algorithm objects over
 * the C ABI (topk.h), so reference-style callers switch by changing a type.
 *
 * Mirrors (avionicscode/TopK):
 *   tuple_<T,Z>                      cpu/AA.h:29-35   (tid, score)
 *   AA<T,Z> get_cdata/get_res/get_thres/benchmark
 *                                    cpu/AA.h:64-124, 194-223
 *   TPAc<T,Z> init/findTopK*         cpu/TPAc.h:15-32
 *   PTA<T,Z>  init/findTopK*         cpu/PTA.h:60-80, 270-278
 *
 * Same conventions as the reference: the data buffer obtained through
 * get_cdata() is column-major (x[m*n+i]), allocated by the caller with
 * malloc (or by the object on first use) and owned by the object, which
 * frees it in its destructor (cpu/AA.h:147-150); init() normalises the store
 * (cpu/TPAc.h:28-32, cpu/PTA.h:271) unless set_normalize(false); get_res()
 * is ascending by score (res[0] is the k-th, cpu/TPAc.h:198-206).  The
 * store is uploaded to the GPU in init(); afterwards the host buffer is no
 * longer read.  Differences: ties rank by lower id; errors throw
 * topk_amd::error (with the topk_status) instead of exit(1).
 */
#ifndef TOPK_AMD_HPP
#define TOPK_AMD_HPP

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "topk.h"

namespace topk_amd {

template <class T, class Z>
struct tuple_ {
    tuple_() : tid(0), score(0) {}
    tuple_(Z t, T s) : tid(t), score(s) {}
    Z tid;
    T score;
};

class error : public std::runtime_error {
public:
    error(int st, const std::string &what)
        : std::runtime_error(what + ": " + topk_strerror(st)), status(st) {}
    int status;
};

inline void check(int st, const char *what) {
    if (st != TOPK_OK) throw error(st, what);
}

/* Base of the GPU algorithm objects (float scores, uint64_t ids). */
class AA {
public:
    AA(uint64_t n, uint64_t d) : n_(n), d_(d) {}
    virtual ~AA() {
        if (ctx_) topk_ctx_destroy(ctx_);
        std::free(cdata_);
    }
    AA(const AA &) = delete;
    AA &operator=(const AA &) = delete;

    float *&get_cdata() {
        if (!cdata_) cdata_ = static_cast<float *>(std::malloc(sizeof(float) * n_ * d_));
        return cdata_;
    }
    void set_cdata(float *c) { cdata_ = c; }
    std::vector<tuple_<float, uint64_t>> get_res() const { return res_; }
    float get_thres() const { return threshold_; }
    std::string get_algo() const { return algo_; }
    void set_normalize(bool on) { normalize_ = on; }
    void set_iter(uint32_t iter) { iter_ = iter ? iter : 1; }
    void reset_clocks() { tt_processing_ = 0; }
    double tt_init() const { return tt_init_; }
    double tt_processing() const { return tt_processing_; }
    uint64_t tuple_count() const { return tuple_count_; }
    topk_ctx *ctx() const { return ctx_; }

    /* Same stdout keys as AA::benchmark (cpu/AA.h:194-223). */
    void benchmark() const {
        std::cout << std::fixed << std::setprecision(8);
        std::cout << "< Benchmark for " << algo_ << " algorithm >" << std::endl;
        std::cout << "tt_init: " << tt_init_ << std::endl;
        std::cout << "tt_procesing: " << tt_processing_ / iter_ << std::endl;
        std::cout << "tuple_count: " << tuple_count_ << std::endl;
    }

protected:
    static double now_ms() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
            .count();
    }
    void upload_and_init(uint32_t flags) {
        const double t0 = now_ms();
        if (!ctx_) {
            if (!cdata_) throw error(TOPK_ERR_INVALID, "init: no data (fill get_cdata() first)");
            check(topk_ctx_create(&ctx_, cdata_, n_, (uint32_t)d_, 0), "topk_ctx_create");
        }
        if (normalize_) flags |= TOPK_INIT_NORMALIZE;
        if (flags) check(topk_init(ctx_, flags), "topk_init");
        tt_init_ = now_ms() - t0;
    }
    template <class F>
    void run(uint64_t k, F &&call) {
        if (!ctx_) throw error(TOPK_ERR_STATE, "findTopK before init()");
        ids_.resize(k);
        scores_.resize(k);
        const double t0 = now_ms();
        float thr = 0;
        check(call(ids_.data(), scores_.data(), &thr), algo_.c_str());
        tt_processing_ += now_ms() - t0;
        const uint64_t c = std::min<uint64_t>(k, n_);
        res_.clear();
        for (uint64_t i = c; i-- > 0;) res_.emplace_back(ids_[i], scores_[i]);  // ascending
        threshold_ = thr;
        topk_stats s;
        if (topk_get_stats(ctx_, &s) == TOPK_OK) tuple_count_ += s.tuple_count;
    }

    uint64_t n_, d_;
    float *cdata_ = nullptr;
    topk_ctx *ctx_ = nullptr;
    bool normalize_ = true;
    std::string algo_;
    std::vector<tuple_<float, uint64_t>> res_;
    std::vector<uint64_t> ids_;
    std::vector<float> scores_;
    float threshold_ = 0;
    double tt_init_ = 0, tt_processing_ = 0;
    uint64_t tuple_count_ = 0;
    uint32_t iter_ = 1;
};

/* Full-table evaluation on the GPU (TPAc, cpu/TPAc.h:15-358). */
class TPAc : public AA {
public:
    TPAc(uint64_t n, uint64_t d) : AA(n, d) { algo_ = "TPAc (gpu fte)"; }
    void init() { upload_and_init(0); }
    void findTopK(uint64_t k, uint8_t qq, const float *weights, const uint8_t *attr) {
        run(k, [&](uint64_t *ids, float *sc, float *thr) {
            return topk_fte(ctx_, (uint32_t)k, qq, weights, attr, ids, sc, thr);
        });
    }
    void findTopKscalar(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
    void findTopKsimd(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
    void findTopKthreads(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
};

/* Partition-pruned evaluation on the GPU (PTA, cpu/PTA.h:60-671). */
class PTA : public AA {
public:
    PTA(uint64_t n, uint64_t d) : AA(n, d) { algo_ = "PTA (gpu)"; }
    void init() { upload_and_init(TOPK_INIT_PTA); }
    void findTopK(uint64_t k, uint8_t qq, const float *weights, const uint8_t *attr) {
        run(k, [&](uint64_t *ids, float *sc, float *thr) {
            return topk_pta(ctx_, (uint32_t)k, qq, weights, attr, ids, sc, thr);
        });
    }
    void findTopKscalar(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
    void findTopKsimd(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
    void findTopKthreads(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
    void findTopKthreads2(uint64_t k, uint8_t qq, const float *w, const uint8_t *a) { findTopK(k, qq, w, a); }
};

/* Many queries per call (findTopKsimdMQ, cpu/TPAc.h:209-254, driven by
 * validation/bench_cpu.h:198-235): add queries, run them as one batch, read
 * each query's ascending result. */
class MultiQuery : public AA {
public:
    MultiQuery(uint64_t n, uint64_t d) : AA(n, d) { algo_ = "MQ (gpu batch)"; }
    void init() { upload_and_init(TOPK_INIT_PTA); }
    void add(uint64_t k, uint8_t qq, const float *weights, const uint8_t *attr) {
        topk_query q{};
        q.k = (uint32_t)k;
        q.qq = qq;
        for (uint32_t j = 0; j < qq && j < TOPK_MAX_DIMS; j++) q.attr[j] = attr[j];
        for (uint32_t a = 0; a < d_ && a < TOPK_MAX_DIMS; a++) q.weights[a] = weights[a];
        qs_.push_back(q);
    }
    void clear() { qs_.clear(); }
    void run_all() {
        if (!ctx_) throw error(TOPK_ERR_STATE, "run_all before init()");
        uint32_t ks = 1;
        for (const auto &q : qs_) ks = std::max(ks, q.k);
        ks_ = ks;
        ids_.resize((size_t)qs_.size() * ks);
        scores_.resize((size_t)qs_.size() * ks);
        thr_.resize(qs_.size());
        const double t0 = now_ms();
        check(topk_batch(ctx_, (uint32_t)qs_.size(), qs_.data(), ks, ids_.data(), scores_.data(), thr_.data()),
              "topk_batch");
        tt_processing_ += now_ms() - t0;
        topk_stats s;
        if (topk_get_stats(ctx_, &s) == TOPK_OK) tuple_count_ += s.tuple_count;
    }
    size_t size() const { return qs_.size(); }
    std::vector<tuple_<float, uint64_t>> get_res(size_t q) const {
        std::vector<tuple_<float, uint64_t>> r;
        const uint64_t c = std::min<uint64_t>(qs_[q].k, n_);
        for (uint64_t i = c; i-- > 0;) r.emplace_back(ids_[q * ks_ + i], scores_[q * ks_ + i]);
        return r;
    }
    float get_thres(size_t q) const { return thr_[q]; }

    /* The multi-query form of AA::benchmark (cpu/AA.h:204-211, IMP >= 3):
     * tt_procesing is the slowest worker's time per iteration (here one
     * batch call per iteration: the call's time; the reference divides that
     * figure by iter once more on this line), tuples_per_second is
     * workload / that time in seconds. */
    void benchmark_mq(uint64_t workload) const {
        const double per_iter = tt_processing_ / iter_;
        std::cout << std::fixed << std::setprecision(8);
        std::cout << "< Benchmark for " << algo_ << " algorithm >" << std::endl;
        std::cout << "tt_init: " << tt_init_ << std::endl;
        std::cout << "tt_procesing: " << per_iter << std::endl;
        std::cout << "tuples_per_second: " << (per_iter > 0 ? workload / (per_iter / 1000.0) : 0.0) << std::endl;
        std::cout << "tuple_count: " << tuple_count_ << std::endl;
    }

private:
    std::vector<topk_query> qs_;
    std::vector<float> thr_;
    uint32_t ks_ = 1;
};

/* Data files with the reference's loader interface (File<T>,
 * input/File.h:14-83, float only): load() parses the text on the GPU
 * (topk_csv_read), gen() runs the synthetic generator, store() writes "%f"
 * rows.  With set_transpose(true) the table is column-major x[m*n + i]
 * (read_scanf_t, :264-295), else row-major (read_scanf, :233-262).  Errors
 * throw topk_amd::error where the reference exits. */
class File {
public:
    File(std::string fname, bool /*gpu*/, uint64_t n, uint64_t d) : fname_(std::move(fname)), n_(n), d_(d) {}
    File(std::string fname, bool /*gpu*/) : fname_(std::move(fname)) {
        uint32_t d = 0;
        check(topk_csv_count(fname_.c_str(), &n_, &d), "topk_csv_count");
        d_ = d;
    }
    uint64_t items() const { return d_; }
    uint64_t rows() const { return n_; }
    void set_transpose(bool t) { transpose_ = t; }
    void set_seed(uint64_t s) { seed_ = s; }
    const topk_ingest_report &report() const { return report_; }

    void load(float *&data) {
        if (!data) data = alloc();
        check(topk_csv_read(fname_.c_str(), n_, (uint32_t)d_, transpose_ ? 0u : TOPK_ROWMAJOR, data, &report_),
              "topk_csv_read");
    }
    void gen(float *&data, int8_t type) {
        if (type < 0 || type > 3)
            throw error(TOPK_ERR_INVALID, "gen: type 0 (correlated), 1 (independent), 2 (anti), 3 (mixed)");
        if (!data) data = alloc();
        if (transpose_) {
            check(topk_generate(data, n_, (uint32_t)d_, n_, type, seed_, 0, n_, 0), "topk_generate");
        } else {
            std::vector<float> c(n_ * d_);
            check(topk_generate(c.data(), n_, (uint32_t)d_, n_, type, seed_, 0, n_, 0), "topk_generate");
            for (uint64_t i = 0; i < n_; i++)
                for (uint64_t m = 0; m < d_; m++) data[i * d_ + m] = c[m * n_ + i];
        }
    }
    void store(const std::string &fname, const float *data) {
        check(topk_csv_write(fname.c_str(), data, n_, (uint32_t)d_, transpose_ ? 0u : TOPK_ROWMAJOR),
              "topk_csv_write");
    }

private:
    float *alloc() const {
        const size_t bytes = ((sizeof(float) * n_ * d_ + 1023) / 1024) * 1024;
        float *p = static_cast<float *>(std::aligned_alloc(1024, bytes));
        if (!p) throw error(TOPK_ERR_NOMEM, "File: host buffer");
        return p;
    }
    std::string fname_;
    uint64_t n_ = 0, d_ = 0;
    bool transpose_ = false;
    uint64_t seed_ = 0;
    topk_ingest_report report_{};
};

}  // namespace topk_amd

#endif
