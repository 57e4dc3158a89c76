// Host-side C++ of libgvs (HNSW builder, index file formats, data generator)
// built with AddressSanitizer and UndefinedBehaviorSanitizer and driven
// through its C ABI: multithreaded and single-thread builds, the NSG-style
// conversion, hnswlib .bin / NSG .nsg / FPGA-layout round trips, the
// reference-written golden indexes, and the error paths for truncated,
// malformed and missing files.  Run by tests/test_sanitizers.py; any
// sanitizer report aborts with a non-zero status.
//
//   host_sanitize <tmp dir> <golden .hnsw.bin> <dim> <metric> [<golden> <dim> <metric> ...]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gvs.h"

static int failures = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "CHECK failed at line %d: %s (%s)\n", __LINE__, #c, \
                         gvs_last_error());                                       \
            failures++;                                                           \
        }                                                                         \
    } while (0)

static void check_graph(const gvs_host_graph *g)
{
    gvs_graph v;
    CHECK(gvs_host_graph_view(g, &v) == GVS_OK);
    CHECK(v.n > 0 && v.r0 > 0 && v.entry >= 0 && v.entry < v.n);
    for (int64_t i = 0; i < v.n * v.r0; i++) CHECK(v.adj0[i] >= -1 && v.adj0[i] < v.n);
    for (int64_t i = 0; v.max_level > 0 && i < v.up_rows * v.ru; i++) CHECK(v.adj_up[i] >= -1 && v.adj_up[i] < v.n);
}

static bool same_graph(const gvs_host_graph *a, const gvs_host_graph *b)
{
    gvs_graph u, v;
    gvs_host_graph_view(a, &u);
    gvs_host_graph_view(b, &v);
    if (u.n != v.n || u.r0 != v.r0 || u.dim != v.dim || u.entry != v.entry || u.max_level != v.max_level) return false;
    if (std::memcmp(u.adj0, v.adj0, sizeof(int32_t) * u.n * u.r0) != 0) return false;
    for (int64_t i = 0; i < u.n; i++)
        if (std::memcmp(u.vectors + i * u.ld_vectors, v.vectors + i * v.ld_vectors, sizeof(float) * u.dim) != 0)
            return false;
    return true;
}

static void write_bytes(const std::string &path, const void *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (n) std::fwrite(p, 1, n, f);
    std::fclose(f);
}

static std::vector<char> read_bytes(const std::string &path)
{
    std::vector<char> v;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return v;
    std::fseek(f, 0, SEEK_END);
    v.resize((size_t)std::ftell(f));
    std::fseek(f, 0, SEEK_SET);
    if (!v.empty() && std::fread(v.data(), 1, v.size(), f) != v.size()) v.clear();
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 2 || (argc - 2) % 3 != 0) {
        std::fprintf(stderr, "usage: host_sanitize <tmp dir> [<golden.bin> <dim> <metric>]...\n");
        return 2;
    }
    const std::string tmp = argv[1];
    const int64_t n = 3000;
    const int dim = 24;
    std::vector<float> x((size_t)n * dim);
    CHECK(gvs_gen_mixture(5, 6, n, dim, 16, 4.0f, 1.0f, 0, x.data(), 4) == GVS_OK);
    for (int metric : {GVS_METRIC_L2, GVS_METRIC_IP}) {
        gvs_host_graph *mt = nullptr, *st = nullptr;
        CHECK(gvs_hnsw_build(x.data(), n, dim, metric, 8, 60, 100, 4, 0, &mt) == GVS_OK);   // 4 threads
        CHECK(gvs_hnsw_build(x.data(), n, dim, metric, 8, 60, 100, 1, 1, &st) == GVS_OK);   // hnswlib order
        if (!mt || !st) return 1;
        check_graph(mt);
        check_graph(st);
        // hnswlib .bin round trip
        const std::string bin = tmp + "/rt.bin";
        CHECK(gvs_hnswlib_save(st, bin.c_str()) == GVS_OK);
        gvs_host_graph *rt = nullptr;
        CHECK(gvs_hnswlib_load(bin.c_str(), dim, metric, &rt) == GVS_OK);
        if (rt) {
            CHECK(same_graph(st, rt));
            gvs_host_graph_free(rt);
        }
        // truncated / malformed copies fail cleanly
        std::vector<char> bytes = read_bytes(bin);
        for (size_t cut : {(size_t)0, (size_t)40, bytes.size() / 2, bytes.size() - 3}) {
            write_bytes(tmp + "/cut.bin", bytes.data(), cut);
            gvs_host_graph *bad = nullptr;
            CHECK(gvs_hnswlib_load((tmp + "/cut.bin").c_str(), dim, metric, &bad) != GVS_OK);
            if (bad) gvs_host_graph_free(bad);
        }
        CHECK(gvs_hnswlib_load(bin.c_str(), dim + 1, metric, &rt) != GVS_OK);  // wrong dim
        // NSG-style conversion and the FPGA-layout directory, 1 and 4 channels
        gvs_host_graph *flat = nullptr;
        CHECK(gvs_nsg_from_hnsw(mt, 16, &flat) == GVS_OK);
        if (flat) check_graph(flat);
        for (int nc : {1, 4}) {
            const std::string d0 = tmp + "/fpga_h" + std::to_string(nc), d1 = tmp + "/fpga_n" + std::to_string(nc);
            std::string mk = "mkdir -p " + d0 + " " + d1;
            CHECK(std::system(mk.c_str()) == 0);
            CHECK(gvs_fpga_save(mt, d0.c_str(), nc) == GVS_OK);
            gvs_host_graph *h = nullptr;
            CHECK(gvs_fpga_load(d0.c_str(), 0, dim, nc, metric, &h) == GVS_OK);
            if (h) {
                CHECK(same_graph(mt, h));
                gvs_host_graph_free(h);
            }
            if (flat) {
                CHECK(gvs_fpga_save(flat, d1.c_str(), nc) == GVS_OK);
                gvs_host_graph *f2 = nullptr;
                CHECK(gvs_fpga_load(d1.c_str(), 1, dim, nc, metric, &f2) == GVS_OK);
                if (f2) {
                    CHECK(same_graph(flat, f2));
                    gvs_host_graph_free(f2);
                }
                gvs_host_graph *bad = nullptr;
                CHECK(gvs_fpga_load(d1.c_str(), 1, dim, nc + 1, metric, &bad) != GVS_OK);  // wrong channel count
                if (bad) gvs_host_graph_free(bad);
                // a link record pointing past n
                const std::string lf = d1 + "/ground_links_" + std::to_string(nc) + "_chan_0.bin";
                std::vector<char> lb = read_bytes(lf);
                if (lb.size() >= 68) {
                    const uint32_t big = 0x7FFFFFFFu;
                    std::memcpy(lb.data() + 64, &big, 4);
                    write_bytes(lf, lb.data(), lb.size());
                    CHECK(gvs_fpga_load(d1.c_str(), 1, dim, nc, metric, &bad) != GVS_OK);
                    if (bad) gvs_host_graph_free(bad);
                }
            }
        }
        // NSG .nsg reader: a valid file, then a bad id and a truncated row
        {
            std::vector<uint32_t> w = {6, 17};
            for (int64_t i = 0; i < n; i++) {
                w.push_back(3);
                for (int j = 0; j < 3; j++) w.push_back((uint32_t)((i * 7 + j * 13 + 1) % n));
            }
            const std::string nf = tmp + "/g.nsg";
            write_bytes(nf, w.data(), w.size() * 4);
            gvs_host_graph *g = nullptr;
            CHECK(gvs_nsg_load(nf.c_str(), x.data(), n, dim, metric, &g) == GVS_OK);
            if (g) {
                check_graph(g);
                gvs_host_graph_free(g);
            }
            w[3] = (uint32_t)n + 5;
            write_bytes(nf, w.data(), w.size() * 4);
            CHECK(gvs_nsg_load(nf.c_str(), x.data(), n, dim, metric, &g) != GVS_OK);
            w[3] = 1;
            write_bytes(nf, w.data(), (w.size() - 2) * 4);
            CHECK(gvs_nsg_load(nf.c_str(), x.data(), n, dim, metric, &g) != GVS_OK);
            w[2] = 1000;  // a row longer than the width
            write_bytes(nf, w.data(), w.size() * 4);
            CHECK(gvs_nsg_load(nf.c_str(), x.data(), n, dim, metric, &g) != GVS_OK);
        }
        if (flat) gvs_host_graph_free(flat);
        gvs_host_graph_free(mt);
        gvs_host_graph_free(st);
    }
    // invalid arguments and missing files
    gvs_host_graph *g = nullptr;
    CHECK(gvs_hnsw_build(x.data(), n, dim, GVS_METRIC_L2, 1, 60, 100, 1, 0, &g) != GVS_OK);
    CHECK(gvs_hnsw_build(nullptr, n, dim, GVS_METRIC_L2, 8, 60, 100, 1, 0, &g) != GVS_OK);
    CHECK(gvs_hnswlib_load((tmp + "/missing.bin").c_str(), dim, GVS_METRIC_L2, &g) != GVS_OK);
    CHECK(gvs_fpga_load((tmp + "/missing").c_str(), 0, dim, 1, GVS_METRIC_L2, &g) != GVS_OK);
    // the reference-written golden indexes
    for (int a = 2; a + 2 < argc; a += 3) {
        gvs_host_graph *r = nullptr;
        CHECK(gvs_hnswlib_load(argv[a], std::atoi(argv[a + 1]), std::atoi(argv[a + 2]), &r) == GVS_OK);
        if (r) {
            check_graph(r);
            const std::string out = tmp + "/golden_rt.bin";
            CHECK(gvs_hnswlib_save(r, out.c_str()) == GVS_OK);
            // same size; the reference leaves stale ids in unused link slots, so
            // equality is of the graph read back, not of every byte
            CHECK(read_bytes(out).size() == read_bytes(argv[a]).size());
            gvs_host_graph *r2 = nullptr;
            CHECK(gvs_hnswlib_load(out.c_str(), std::atoi(argv[a + 1]), std::atoi(argv[a + 2]), &r2) == GVS_OK);
            if (r2) {
                CHECK(same_graph(r, r2));
                gvs_host_graph_free(r2);
            }
            gvs_host_graph_free(r);
        }
    }
    std::printf("host_sanitize: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
