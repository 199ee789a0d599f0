// Reference-style benchmark driver over the C++ mirror (include/topk_amd.hpp):
// the loops and stdout keys of validation/bench_cpu.h (bench_tpac :155-237,
// bench_pta :323-402), on the GPU engine.
//
//   bench_topk [-n N] [-d D] [-s KS] [-e KE] [-q QD] [-i ITER] [-t DISTR] [-r SEED]
//              [-a fte|pta|mq|mq4] [-p IMP] [-m QM] [-w WORKLOAD] [-f TEXTFILE]
//              [--no-normalize] [--check] [--mq-thresholds]
//
// The compile-time knobs of the reference's Makefile (:18-35) are flags:
//   -a  algorithm object: fte (TPAc), pta (PTA); mq and mq4 are shorthands
//       for -p 3 and -p 4
//   -p  IMP: 0 scalar, 1 SIMD, 2 threads (one GPU call each), 3 random-
//       attribute multiple queries (bench_cpu.h:198-215: WORKLOAD queries,
//       each on a random qq in [2, d], one report per k), 4 same-attribute
//       multiple queries (:216-235: WORKLOAD queries on the same qq, one
//       report per (k, qq))
//   -m  QM: 0 = the last qq attributes ascending, 1 = the first qq
//       (the attr[][] tables, bench_cpu.h:25-68)
// k doubles from KS to KE; qq runs 2..d in steps of QD; weights 1.0 (the
// Q0 vector, bench_cpu.h:19); one warm-up call, then ITER timed calls.
// Multiple-query modes run ITER batch calls of the WORKLOAD queries.
// Each per-query report ends with "threshold: " (9 significant digits: the
// binary32 value round-trips); --mq-thresholds prints one
// "mq_threshold: j qq k value" line per query of the last batch.  --check
// runs TPAc and PTA on the same store and compares ids and thresholds like
// AA::compare / compare_t (cpu/AA.h:155-189), printing PASSED or FAILED.
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <string>

#include "topk_amd.hpp"

namespace {

struct Args {
    uint64_t n = 1 << 20;
    uint32_t d = 8;
    uint64_t ks = 16, ke = 16;
    uint32_t qd = 1;
    uint32_t iter = 10;
    int distr = 1;
    uint64_t seed = 0;
    std::string algo = "fte";
    int imp = 2;
    int qm = 0;
    uint32_t workload = 1024;
    bool normalize = true;
    bool check = false;
    bool mq_thresholds = false;
    std::string file;  // load this text file instead of generating (the reference's LD != 1)
};

bool parse(int argc, char **argv, Args &a) {
    for (int i = 1; i < argc; i++) {
        const std::string s = argv[i];
        auto next = [&](uint64_t &v) {
            if (i + 1 >= argc) return false;
            v = std::strtoull(argv[++i], nullptr, 10);
            return true;
        };
        uint64_t v = 0;
        if (s == "-n" && next(v)) a.n = v;
        else if (s == "-d" && next(v)) a.d = (uint32_t)v;
        else if (s == "-s" && next(v)) a.ks = v;
        else if (s == "-e" && next(v)) a.ke = v;
        else if (s == "-q" && next(v)) a.qd = (uint32_t)(v ? v : 1);
        else if (s == "-i" && next(v)) a.iter = (uint32_t)(v ? v : 1);
        else if (s == "-t" && next(v)) a.distr = (int)v;
        else if (s == "-r" && next(v)) a.seed = v;
        else if (s == "-w" && next(v)) a.workload = (uint32_t)v;
        else if (s == "-p" && next(v)) a.imp = (int)v;
        else if (s == "-m" && next(v)) a.qm = (int)v;
        else if (s == "-a" && i + 1 < argc) a.algo = argv[++i];
        else if (s == "-f" && i + 1 < argc) a.file = argv[++i];
        else if (s == "--no-normalize") a.normalize = false;
        else if (s == "--check") a.check = true;
        else if (s == "--mq-thresholds") a.mq_thresholds = true;
        else return false;
    }
    if (a.algo == "mq") a.imp = 3;
    if (a.algo == "mq4") a.imp = 4;
    return a.d >= 2 && a.d <= TOPK_MAX_DIMS && a.n > 0 && a.ks >= 1 && a.ke >= a.ks && a.imp >= 0 &&
           a.imp <= 4 && (a.qm == 0 || a.qm == 1) && a.workload >= 1 &&
           (a.algo == "fte" || a.algo == "pta" || a.algo == "mq" || a.algo == "mq4");
}

template <class Algo>
void fill(Algo &al, const Args &a) {
    topk_amd::File f(a.file, false, a.n, a.d);
    f.set_transpose(true);
    f.set_seed(a.seed);
    if (!a.file.empty()) {
        std::cout << "Loading data from file !!!" << std::endl;
        f.load(al.get_cdata());
    } else {
        f.gen(al.get_cdata(), (int8_t)a.distr);
    }
    al.set_normalize(a.normalize);
}

// Row qq - 2 of the reference's attr[][] table: QM = 0 lists the last qq
// attributes ascending, QM = 1 the first qq (bench_cpu.h:25-68).
std::vector<uint8_t> query_attrs(uint32_t d, uint32_t qq, int qm) {
    std::vector<uint8_t> at(qq);
    for (uint32_t j = 0; j < qq; j++) at[j] = (uint8_t)(qm == 0 ? d - qq + j : j);
    return at;
}

void print_threshold(const char *key, float t) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.9g", (double)t);
    std::cout << key << buf << std::endl;
}

void header(uint64_t n, uint32_t x, uint64_t k) {
    std::cout << "Benchmark <<<-------------" << n << "," << x << "," << k << "------------->>> " << std::endl;
}

// IMP 0..2: one query at a time.
template <class Algo>
void sweep(Algo &al, const Args &a, const std::vector<float> &w) {
    auto call = [&](uint64_t k, uint32_t qq, const uint8_t *at) {
        if (a.imp == 0) al.findTopKscalar(k, (uint8_t)qq, w.data(), at);
        else if (a.imp == 1) al.findTopKsimd(k, (uint8_t)qq, w.data(), at);
        else al.findTopKthreads(k, (uint8_t)qq, w.data(), at);
    };
    for (uint64_t k = a.ks; k <= a.ke; k *= 2) {
        for (uint32_t qq = 2; qq <= a.d; qq += a.qd) {
            const auto at = query_attrs(a.d, qq, a.qm);
            header(a.n, qq, k);
            call(k, qq, at.data());  // warm-up
            al.reset_clocks();
            al.set_iter(a.iter);
            for (uint32_t it = 0; it < a.iter; it++) call(k, qq, at.data());
            al.benchmark();
            print_threshold("threshold: ", al.get_thres());
        }
    }
}

// One timed multiple-query measurement: a warm-up batch, then ITER batches.
void run_mq(topk_amd::MultiQuery &mq, const Args &a, const std::vector<uint32_t> &qqs, uint64_t k) {
    mq.run_all();  // warm-up
    mq.reset_clocks();
    mq.set_iter(a.iter);
    for (uint32_t it = 0; it < a.iter; it++) mq.run_all();
    mq.benchmark_mq(a.workload);
    if (a.mq_thresholds) {
        for (size_t j = 0; j < mq.size(); j++) {
            char buf[96];
            std::snprintf(buf, sizeof(buf), "mq_threshold: %zu %u %llu %.9g", j, qqs[j], (unsigned long long)k,
                          (double)mq.get_thres(j));
            std::cout << buf << std::endl;
        }
    }
}

// IMP 3 (random attribute counts, bench_cpu.h:198-215) and IMP 4 (the same
// attributes for every query of the workload, :216-235).
void multi_query(const Args &a, const std::vector<float> &w) {
    topk_amd::MultiQuery mq(a.n, a.d);
    fill(mq, a);
    mq.init();
    std::vector<std::vector<uint8_t>> keep;  // attr rows outlive the batch
    if (a.imp == 3) {
        // random_workload (bench_cpu.h:61-67) draws rand() % (d - 1) + 2 with a
        // time seed; here a seeded generator, one draw per query
        std::mt19937_64 rng(a.seed + 1);
        std::vector<uint32_t> qqs(a.workload);
        for (auto &q : qqs) q = 2 + (uint32_t)(rng() % (a.d - 1));
        std::cout << "<<<Random Attribute Multiple Queries - (gpu batch) >>>" << std::endl;
        for (uint64_t k = a.ks; k <= a.ke; k *= 2) {
            mq.clear();
            keep.clear();
            for (uint32_t j = 0; j < a.workload; j++) {
                keep.push_back(query_attrs(a.d, qqs[j], a.qm));
                mq.add(k, (uint8_t)qqs[j], w.data(), keep.back().data());
            }
            header(a.n, a.d, k);
            run_mq(mq, a, qqs, k);
        }
    } else {
        std::cout << "<<<Same Attribute Multiple Queries - (gpu batch) >>>" << std::endl;
        for (uint64_t k = a.ks; k <= a.ke; k *= 2) {
            for (uint32_t qq = 2; qq <= a.d; qq += a.qd) {
                mq.clear();
                keep.assign(1, query_attrs(a.d, qq, a.qm));
                for (uint32_t j = 0; j < a.workload; j++) mq.add(k, (uint8_t)qq, w.data(), keep[0].data());
                header(a.n, qq, k);
                run_mq(mq, a, std::vector<uint32_t>(a.workload, qq), k);
                mq.reset_clocks();
            }
        }
    }
}

int compare(const topk_amd::AA &x, const topk_amd::AA &y) {
    std::map<uint64_t, float> m;
    for (const auto &t : x.get_res()) m[t.tid] = t.score;
    bool ok = x.get_res().size() == y.get_res().size();
    for (const auto &t : y.get_res()) ok = ok && m.count(t.tid) && m[t.tid] == t.score;
    ok = ok && x.get_thres() == y.get_thres();
    std::cout << "(" << x.get_algo() << ") compare result to (" << y.get_algo() << ") ( "
              << (ok ? "PASSED" : "FAILED") << " )" << std::endl;
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    Args a;
    if (!parse(argc, argv, a)) {
        std::fprintf(stderr,
                     "usage: bench_topk [-n N] [-d D] [-s KS] [-e KE] [-q QD] [-i ITER] [-t DISTR] [-r SEED] "
                     "[-a fte|pta|mq|mq4] [-p IMP 0..4] [-m QM 0|1] [-w WORKLOAD] [-f TEXTFILE] "
                     "[--no-normalize] [--check] [--mq-thresholds]\n");
        return 2;
    }
    const std::vector<float> w(a.d, 1.0f);
    try {
        if (a.check) {
            topk_amd::TPAc t(a.n, a.d);
            topk_amd::PTA p(a.n, a.d);
            fill(t, a);
            fill(p, a);
            t.init();
            p.init();
            int bad = 0;
            for (uint64_t k = a.ks; k <= a.ke; k *= 2)
                for (uint32_t qq = 2; qq <= a.d; qq += a.qd) {
                    const auto at = query_attrs(a.d, qq, a.qm);
                    t.findTopK(k, (uint8_t)qq, w.data(), at.data());
                    p.findTopK(k, (uint8_t)qq, w.data(), at.data());
                    bad |= compare(t, p);
                }
            return bad;
        }
        if (a.imp >= 3) {
            multi_query(a, w);
        } else if (a.algo == "fte") {
            topk_amd::TPAc t(a.n, a.d);
            fill(t, a);
            t.init();
            sweep(t, a, w);
        } else {
            topk_amd::PTA p(a.n, a.d);
            fill(p, a);
            p.init();
            sweep(p, a, w);
        }
    } catch (const topk_amd::error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
