// gvs_search: C++ host program over libgvs.so (through the hnswlib-style
// wrapper in include/gvs_hnswlib.hpp).  It takes the place of the FPGA host
// (FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src/host.cpp):
// load an index, run a query batch with ef / mc / mg, report Recall@1 and
// Recall@10 against ground truth the way host.cpp:1025-1070 does, and the
// average layer-0 hops and evaluations (its mem_debug report).
//
//   gvs_search INDEX.bin DIM l2|ip QUERIES.fvecs K EF [--mc N] [--mg N]
//              [--gt GT.ivecs] [--out PREFIX] [--storage f32|bf16|fp8] [--rerank]
//
// .fvecs / .ivecs: per row an int32 count followed by that many float32 /
// int32 values (the SIFT/Deep file format the reference's datasets use).
// --out PREFIX writes PREFIX.ids (int64, nq x K) and PREFIX.dists (float32).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "gvs_hnswlib.hpp"

template <typename T>
static bool read_vecs(const std::string &path, std::vector<T> &out, int &dim, size_t &rows)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.clear();
    rows = 0;
    dim = -1;
    int32_t d;
    while (f.read((char *)&d, 4)) {
        if (dim < 0) dim = d;
        if (d != dim || d <= 0) return false;
        size_t off = out.size();
        out.resize(off + (size_t)d);
        if (!f.read((char *)(out.data() + off), sizeof(T) * (size_t)d)) return false;
        rows++;
    }
    return dim > 0;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::fprintf(stderr,
                     "usage: %s INDEX.bin DIM l2|ip QUERIES.fvecs K EF [--mc N] [--mg N] [--gt GT.ivecs] "
                     "[--out PREFIX] [--storage f32|bf16|fp8] [--rerank]\n",
                     argv[0]);
        return 2;
    }
    const std::string index_path = argv[1];
    const int dim = std::atoi(argv[2]);
    const int metric = std::strcmp(argv[3], "ip") == 0 ? GVS_METRIC_IP : GVS_METRIC_L2;
    const std::string qpath = argv[4];
    const int k = std::atoi(argv[5]);
    const int ef = std::atoi(argv[6]);
    int mc = 1, mg = 1, storage = GVS_STORE_F32;
    bool rerank = false;
    std::string gt_path, out_prefix;
    for (int i = 7; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--mc" && i + 1 < argc) mc = std::atoi(argv[++i]);
        else if (a == "--mg" && i + 1 < argc) mg = std::atoi(argv[++i]);
        else if (a == "--gt" && i + 1 < argc) gt_path = argv[++i];
        else if (a == "--out" && i + 1 < argc) out_prefix = argv[++i];
        else if (a == "--storage" && i + 1 < argc) {
            const std::string s = argv[++i];
            storage = s == "bf16" ? GVS_STORE_BF16 : s == "fp8" ? GVS_STORE_FP8 : GVS_STORE_F32;
        } else if (a == "--rerank") rerank = true;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    std::vector<float> q;
    int qdim = 0;
    size_t nq = 0;
    if (!read_vecs(qpath, q, qdim, nq) || qdim != dim) {
        std::fprintf(stderr, "cannot read %s as %d-dim fvecs\n", qpath.c_str(), dim);
        return 1;
    }
    try {
        gvs_hnswlib::HierarchicalNSW index(index_path, dim, metric,
                                          storage == GVS_STORE_F32 ? 0u : (1u << storage));
        index.setEf((size_t)ef);
        index.setDst(mc, mg);
        index.setStorage(storage, rerank);
        std::vector<int64_t> ids(nq * (size_t)k);
        std::vector<float> dists(nq * (size_t)k);
        std::vector<int32_t> stats(nq * GVS_STATS_PER_QUERY);
        index.searchKnnBatch(q.data(), 1, (size_t)k, ids.data(), dists.data());  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        index.searchKnnBatch(q.data(), nq, (size_t)k, ids.data(), dists.data(), stats.data());
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << std::setprecision(10);
        std::cout << "queries=" << nq << " k=" << k << " ef=" << ef << " mc=" << mc << " mg=" << mg << "\n";
        std::cout << "time_ms=" << s * 1e3 << " QPS(host buffers)=" << (double)nq / s << "\n";
        double hops = 0, evals = 0;
        for (size_t i = 0; i < nq; i++) {
            hops += stats[i * GVS_STATS_PER_QUERY + 0];
            evals += stats[i * GVS_STATS_PER_QUERY + 1];
        }
        std::cout << "avg_hops=" << hops / (double)nq << " avg_evaluations=" << evals / (double)nq << "\n";
        if (!gt_path.empty()) {
            std::vector<int32_t> gt;
            int gk = 0;
            size_t gn = 0;
            if (!read_vecs(gt_path, gt, gk, gn) || gn != nq || gk < 1) {
                std::fprintf(stderr, "cannot read %s as ground truth for %zu queries\n", gt_path.c_str(), nq);
                return 1;
            }
            size_t top1 = 0, top10 = 0;
            const int kk = std::min(std::min(10, k), gk);
            for (size_t i = 0; i < nq; i++) {
                if (ids[i * k] == gt[i * gk]) top1++;
                for (int a = 0; a < kk; a++)
                    for (int b = 0; b < kk; b++)
                        if (ids[i * k + b] == gt[i * gk + a]) {
                            top10++;
                            break;
                        }
            }
            std::cout << "Recall@1=" << (double)top1 / (double)nq << "\n";
            std::cout << "Recall@" << kk << "=" << (double)top10 / (double)(nq * kk) << "\n";
        }
        if (!out_prefix.empty()) {
            std::ofstream fi(out_prefix + ".ids", std::ios::binary), fd(out_prefix + ".dists", std::ios::binary);
            fi.write((const char *)ids.data(), sizeof(int64_t) * ids.size());
            fd.write((const char *)dists.data(), sizeof(float) * dists.size());
            if (!fi || !fd) {
                std::fprintf(stderr, "cannot write %s.*\n", out_prefix.c_str());
                return 1;
            }
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gvs_search: %s\n", e.what());
        return 1;
    }
    return 0;
}
