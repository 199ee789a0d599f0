// Reference-style benchmark driver over the C++ mirror (include/topk_amd.hpp):
// the loops and stdout keys of validation/bench_cpu.h (bench_tpac :155-237,
// bench_pta :323-402, multi-query IMP=3 :198-215), on the GPU engine.
//
//   bench_topk [-n N] [-d D] [-s KS] [-e KE] [-q QD] [-i ITER] [-t DISTR] [-r SEED]
//              [-a fte|pta|mq] [-w WORKLOAD] [--no-normalize] [--check]
//
// k doubles from KS to KE; qq runs 2..d in steps of QD over the last qq
// attributes (the QM=0 tables, bench_cpu.h:25-49) with weights 1.0; one
// warm-up call, then ITER timed calls.  --check runs TPAc and PTA on the same
// store and compares ids and thresholds like AA::compare / compare_t
// (cpu/AA.h:155-189), printing PASSED or FAILED.
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
    uint32_t workload = 1024;
    bool normalize = true;
    bool check = false;
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
        else if (s == "-a" && i + 1 < argc) a.algo = argv[++i];
        else if (s == "-f" && i + 1 < argc) a.file = argv[++i];
        else if (s == "--no-normalize") a.normalize = false;
        else if (s == "--check") a.check = true;
        else return false;
    }
    return a.d >= 2 && a.d <= TOPK_MAX_DIMS && a.n > 0 && a.ks >= 1 && a.ke >= a.ks;
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

std::vector<uint8_t> last_attrs(uint32_t d, uint32_t qq) {
    std::vector<uint8_t> at(qq);
    for (uint32_t j = 0; j < qq; j++) at[j] = (uint8_t)(d - qq + j);
    return at;
}

template <class Algo>
void sweep(Algo &al, const Args &a, const std::vector<float> &w) {
    for (uint64_t k = a.ks; k <= a.ke; k *= 2) {
        for (uint32_t qq = 2; qq <= a.d; qq += a.qd) {
            const auto at = last_attrs(a.d, qq);
            std::cout << "Benchmark <<<-------------" << a.n << "," << qq << "," << k << "------------->>> "
                      << std::endl;
            al.findTopK(k, (uint8_t)qq, w.data(), at.data());  // warm-up
            al.reset_clocks();
            al.set_iter(a.iter);
            for (uint32_t it = 0; it < a.iter; it++) al.findTopK(k, (uint8_t)qq, w.data(), at.data());
            al.benchmark();
            std::cout << "threshold: " << al.get_thres() << std::endl;
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
                     "[-a fte|pta|mq] [-w WORKLOAD] [-f TEXTFILE] [--no-normalize] [--check]\n");
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
                    const auto at = last_attrs(a.d, qq);
                    t.findTopK(k, (uint8_t)qq, w.data(), at.data());
                    p.findTopK(k, (uint8_t)qq, w.data(), at.data());
                    bad |= compare(t, p);
                }
            return bad;
        }
        if (a.algo == "fte") {
            topk_amd::TPAc t(a.n, a.d);
            fill(t, a);
            t.init();
            sweep(t, a, w);
        } else if (a.algo == "pta") {
            topk_amd::PTA p(a.n, a.d);
            fill(p, a);
            p.init();
            sweep(p, a, w);
        } else if (a.algo == "mq") {
            // random qq in [2, d] per query (bench_cpu.h:61-67 random_workload), seeded
            topk_amd::MultiQuery mq(a.n, a.d);
            fill(mq, a);
            mq.init();
            std::mt19937_64 rng(a.seed + 1);
            for (uint64_t k = a.ks; k <= a.ke; k *= 2) {
                mq.clear();
                std::vector<std::vector<uint8_t>> keep;
                for (uint32_t j = 0; j < a.workload; j++) {
                    const uint32_t qq = 2 + (uint32_t)(rng() % (a.d - 1));
                    keep.push_back(last_attrs(a.d, qq));
                    mq.add(k, (uint8_t)qq, w.data(), keep.back().data());
                }
                std::cout << "Benchmark <<<-------------" << a.n << "," << a.d << "," << k
                          << "------------->>> " << std::endl;
                mq.run_all();  // warm-up
                mq.reset_clocks();
                mq.set_iter(a.iter);
                for (uint32_t it = 0; it < a.iter; it++) mq.run_all();
                mq.benchmark();
                std::cout << "tuples_per_second: " << a.workload / (mq.tt_processing() / a.iter / 1000.0)
                          << std::endl;
            }
        } else {
            std::fprintf(stderr, "unknown algorithm %s\n", a.algo.c_str());
            return 2;
        }
    } catch (const topk_amd::error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
