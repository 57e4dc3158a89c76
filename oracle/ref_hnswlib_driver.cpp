// Driver for the reference's vendored hnswlib (header-only), compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`).  The
// binaries go to oracle/_ref/ and are used only by tests/ and
// tests/golden/make_golden.py to pin the oracle, and by bench.py's CPU
// baseline; they are never shipped.  ref_hnsw is the scalar build
// (-DNO_MANUAL_VECTORIZATION, the fixtures' order), ref_hnsw_simd the SIMD
// build (hnswlib's own SSE/AVX distance kernels) with OpenMP.
//
//   ref_hnsw build  <vecs.f32> <N> <D> <l2|ip> <M> <efC> <seed> <out_index.bin>
//   ref_hnsw search <index.bin> <D> <l2|ip> <queries.f32> <nq> <k> <ef> <out_ids.i32> <out_dists.f32>
//   ref_hnsw_simd bench <index.bin> <D> <l2|ip> <queries.f32> <nq> <k> <ef> <threads> <seconds> <out_ids.i32>
//
// bench: searchKnn over the queries with OpenMP across queries on `threads`
// threads, repeated until `seconds` have passed (at least once), then 500
// queries one at a time on one thread (batch=1 latency, first and last
// dropped as perf_test_scripts/run_all_hnsw_search.py:41-44 does); prints one
// JSON line and writes the first pass's ids.
//
// build: HierarchicalNSW(space, N, M, efC, seed) then addPoint(x_i, i) for
//        i = 0..N-1 on one thread (networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:27-66, :982-1112),
//        then saveIndex (hnswalg.h:584-616).
// search: loadIndex, setEf(ef), searchKnnCloserFirst(q, k) (hnswlib.h:171-185).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "hnswlib/hnswlib.h"

static std::vector<float> read_f32(const char *path, size_t count)
{
    std::vector<float> v(count);
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    if (std::fread(v.data(), sizeof(float), count, f) != count) { std::fprintf(stderr, "short read %s\n", path); std::exit(2); }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: see header\n"); return 2; }
    std::string mode = argv[1];
    try {
        if (mode == "build" && argc == 10) {
            size_t n = std::strtoull(argv[3], nullptr, 10);
            size_t d = std::strtoull(argv[4], nullptr, 10);
            std::string metric = argv[5];
            size_t M = std::strtoull(argv[6], nullptr, 10);
            size_t efc = std::strtoull(argv[7], nullptr, 10);
            size_t seed = std::strtoull(argv[8], nullptr, 10);
            std::vector<float> x = read_f32(argv[2], n * d);
            hnswlib::L2Space l2(d);
            hnswlib::InnerProductSpace ip(d);
            hnswlib::SpaceInterface<float> *sp = metric == "ip" ? (hnswlib::SpaceInterface<float> *)&ip
                                                                : (hnswlib::SpaceInterface<float> *)&l2;
            hnswlib::HierarchicalNSW<float> index(sp, n, M, efc, seed);
            for (size_t i = 0; i < n; i++) index.addPoint(x.data() + i * d, i);
            index.saveIndex(argv[9]);
            return 0;
        }
        if (mode == "search" && argc == 11) {
            size_t d = std::strtoull(argv[3], nullptr, 10);
            std::string metric = argv[4];
            size_t nq = std::strtoull(argv[6], nullptr, 10);
            size_t k = std::strtoull(argv[7], nullptr, 10);
            size_t ef = std::strtoull(argv[8], nullptr, 10);
            std::vector<float> q = read_f32(argv[5], nq * d);
            hnswlib::L2Space l2(d);
            hnswlib::InnerProductSpace ip(d);
            hnswlib::SpaceInterface<float> *sp = metric == "ip" ? (hnswlib::SpaceInterface<float> *)&ip
                                                                : (hnswlib::SpaceInterface<float> *)&l2;
            hnswlib::HierarchicalNSW<float> index(sp, argv[2]);
            index.setEf(ef);
            std::vector<int> ids(nq * k, -1);
            std::vector<float> ds(nq * k, INFINITY);
            for (size_t i = 0; i < nq; i++) {
                auto r = index.searchKnnCloserFirst(q.data() + i * d, k);
                for (size_t j = 0; j < r.size(); j++) {
                    ids[i * k + j] = (int)r[j].second;
                    ds[i * k + j] = r[j].first;
                }
            }
            FILE *f = std::fopen(argv[9], "wb");
            std::fwrite(ids.data(), sizeof(int), ids.size(), f);
            std::fclose(f);
            f = std::fopen(argv[10], "wb");
            std::fwrite(ds.data(), sizeof(float), ds.size(), f);
            std::fclose(f);
            return 0;
        }
        if (mode == "bench" && argc == 12) {
            size_t d = std::strtoull(argv[3], nullptr, 10);
            std::string metric = argv[4];
            long nq = std::strtol(argv[6], nullptr, 10);
            size_t k = std::strtoull(argv[7], nullptr, 10);
            size_t ef = std::strtoull(argv[8], nullptr, 10);
            int threads = std::atoi(argv[9]);
            double seconds = std::atof(argv[10]);
            std::vector<float> q = read_f32(argv[5], (size_t)nq * d);
            hnswlib::L2Space l2(d);
            hnswlib::InnerProductSpace ip(d);
            hnswlib::SpaceInterface<float> *sp = metric == "ip" ? (hnswlib::SpaceInterface<float> *)&ip
                                                                : (hnswlib::SpaceInterface<float> *)&l2;
            hnswlib::HierarchicalNSW<float> index(sp, argv[2]);
            index.setEf(ef);
            std::vector<int> ids((size_t)nq * k, -1);
            using clk = std::chrono::steady_clock;
            auto pass = [&](bool keep) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
                for (long i = 0; i < nq; i++) {
                    auto r = index.searchKnnCloserFirst(q.data() + (size_t)i * d, k);
                    if (keep)
                        for (size_t j = 0; j < r.size(); j++) ids[(size_t)i * k + j] = (int)r[j].second;
                }
            };
            pass(false);  // warm-up (page in the index, visited lists)
            const auto t0 = clk::now();
            int passes = 0;
            double el = 0.0;
            do {
                pass(passes == 0);
                passes++;
                el = std::chrono::duration<double>(clk::now() - t0).count();
            } while (el < seconds);
            std::vector<double> lat;
            for (long i = 0; i < std::min<long>(nq, 502); i++) {
                const auto a = clk::now();
                index.searchKnnCloserFirst(q.data() + (size_t)i * d, k);
                lat.push_back(std::chrono::duration<double, std::micro>(clk::now() - a).count());
            }
            if (lat.size() > 2) lat = std::vector<double>(lat.begin() + 1, lat.end() - 1);
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[(size_t)std::min<double>(lat.size() - 1, p * lat.size())]; };
            std::printf("{\"qps\": %.1f, \"passes\": %d, \"queries\": %ld, \"threads\": %d, \"seconds\": %.3f, "
                        "\"batch1_p50_us\": %.1f, \"batch1_p99_us\": %.1f, \"batch1_queries\": %zu}\n",
                        passes * (double)nq / el, passes, nq, threads, el, pct(0.5), pct(0.99), lat.size());
            FILE *f = std::fopen(argv[11], "wb");
            std::fwrite(ids.data(), sizeof(int), ids.size(), f);
            std::fclose(f);
            return 0;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
