// On-disk formats and host-graph plumbing.
//   hnswlib .bin  : networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:584-700
//                   (96-byte header, level-0 blocks of count + maxM0 ids +
//                   vector + size_t label, then per node a u32 byte count and
//                   its upper-level lists); parsed the same way by
//                   vector_search_baselines/scripts_hnsw/hnsw.py:139-262.
//   NSG .nsg      : vector_search_baselines/scripts_nsg/nsg_to_FPGA.py:31-47
//                   (u32 width, u32 ep, then per node u32 k and k u32 ids).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "host_graph.hpp"

namespace gvs {
static thread_local char g_err[512] = "";
void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_err; }
}  // namespace gvs

extern "C" const char *gvs_last_error(void) { return gvs::last_error(); }

namespace {
template <class T>
bool rd(std::ifstream &f, T &v)
{
    f.read(reinterpret_cast<char *>(&v), sizeof(T));
    return (bool)f;
}
template <class T>
void wr(std::ofstream &f, const T &v)
{
    f.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
}  // namespace

extern "C" int gvs_hnswlib_load(const char *path, int32_t dim, int32_t metric, gvs_host_graph **out)
{
    if (!path || dim <= 0 || !out) {
        gvs::set_error("gvs_hnswlib_load: invalid argument");
        return GVS_ERR_ARG;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        gvs::set_error("gvs_hnswlib_load: cannot open %s", path);
        return GVS_ERR_IO;
    }
    uint64_t offsetLevel0, max_elements, count, size_per, label_off, data_off, maxM, maxM0, M, efc;
    int32_t maxlevel;
    uint32_t ep;
    double mult;
    bool ok = rd(f, offsetLevel0) && rd(f, max_elements) && rd(f, count) && rd(f, size_per) &&
              rd(f, label_off) && rd(f, data_off) && rd(f, maxlevel) && rd(f, ep) && rd(f, maxM) &&
              rd(f, maxM0) && rd(f, M) && rd(f, mult) && rd(f, efc);
    if (!ok || offsetLevel0 != 0 || size_per != 4 + 4 * maxM0 + 4 * (uint64_t)dim + 8 ||
        data_off != 4 + 4 * maxM0 || label_off != data_off + 4 * (uint64_t)dim || count == 0 ||
        count > (uint64_t)INT32_MAX || maxM0 > 4096 || maxM > 4096) {
        gvs::set_error("gvs_hnswlib_load: bad header in %s (dim %d?)", path, dim);
        return GVS_ERR_IO;
    }
    gvs_host_graph *g = new gvs_host_graph();
    g->n = (int64_t)count;
    g->dim = dim;
    g->metric = metric;
    g->r0 = (int32_t)maxM0;
    g->ru = (int32_t)maxM;
    g->max_level = maxlevel;
    g->entry = (int32_t)ep;
    g->M = (int32_t)M;
    g->ef_construction = (int32_t)efc;
    g->mult = mult;
    g->vectors.resize((size_t)count * dim);
    g->adj0.assign((size_t)count * maxM0, -1);
    g->labels.resize((size_t)count);
    std::vector<char> blk((size_t)size_per);
    for (uint64_t i = 0; i < count; i++) {
        if (!f.read(blk.data(), (std::streamsize)size_per)) {
            delete g;
            gvs::set_error("gvs_hnswlib_load: truncated level-0 data");
            return GVS_ERR_IO;
        }
        uint16_t c;
        std::memcpy(&c, blk.data(), 2);
        if (c > maxM0) {
            delete g;
            gvs::set_error("gvs_hnswlib_load: bad list count");
            return GVS_ERR_IO;
        }
        for (uint16_t j = 0; j < c; j++) std::memcpy(&g->adj0[i * maxM0 + j], blk.data() + 4 + 4 * j, 4);
        std::memcpy(&g->vectors[i * dim], blk.data() + data_off, 4 * (size_t)dim);
        uint64_t lab;
        std::memcpy(&lab, blk.data() + label_off, 8);
        g->labels[i] = (int64_t)lab;
    }
    g->levels.assign((size_t)count, 0);
    g->upper_off.assign((size_t)count, -1);
    const uint64_t per_level = 4 + 4 * maxM;
    int64_t rows = 0;
    std::vector<std::vector<char>> ul((size_t)count);
    for (uint64_t i = 0; i < count; i++) {
        uint32_t sz;
        if (!rd(f, sz) || sz % per_level != 0) {
            delete g;
            gvs::set_error("gvs_hnswlib_load: bad link list size");
            return GVS_ERR_IO;
        }
        if (sz) {
            ul[i].resize(sz);
            if (!f.read(ul[i].data(), sz)) {
                delete g;
                gvs::set_error("gvs_hnswlib_load: truncated link lists");
                return GVS_ERR_IO;
            }
            g->levels[i] = (int32_t)(sz / per_level);
            g->upper_off[i] = rows;
            rows += g->levels[i];
        }
    }
    g->adj_up.assign((size_t)rows * maxM, -1);
    for (uint64_t i = 0; i < count; i++) {
        for (int32_t l = 0; l < g->levels[i]; l++) {
            const char *src = ul[i].data() + l * per_level;
            uint16_t c;
            std::memcpy(&c, src, 2);
            if (c > maxM) c = (uint16_t)maxM;
            for (uint16_t j = 0; j < c; j++)
                std::memcpy(&g->adj_up[(size_t)(g->upper_off[i] + l) * maxM + j], src + 4 + 4 * j, 4);
        }
    }
    *out = g;
    return GVS_OK;
}

extern "C" int gvs_hnswlib_save(const gvs_host_graph *g, const char *path)
{
    if (!g || !path || g->max_level < 0) {
        gvs::set_error("gvs_hnswlib_save: invalid argument");
        return GVS_ERR_ARG;
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        gvs::set_error("gvs_hnswlib_save: cannot open %s", path);
        return GVS_ERR_IO;
    }
    const uint64_t maxM0 = (uint64_t)g->r0, maxM = (uint64_t)(g->ru > 0 ? g->ru : g->r0 / 2);
    const uint64_t data_off = 4 + 4 * maxM0, label_off = data_off + 4 * (uint64_t)g->dim;
    const uint64_t size_per = label_off + 8;
    wr(f, (uint64_t)0);
    wr(f, (uint64_t)g->n);
    wr(f, (uint64_t)g->n);
    wr(f, size_per);
    wr(f, label_off);
    wr(f, data_off);
    wr(f, (int32_t)g->max_level);
    wr(f, (uint32_t)g->entry);
    wr(f, maxM);
    wr(f, maxM0);
    wr(f, (uint64_t)(g->M > 0 ? g->M : (int32_t)maxM));
    wr(f, g->mult);
    wr(f, (uint64_t)g->ef_construction);
    std::vector<char> blk((size_t)size_per);
    for (int64_t i = 0; i < g->n; i++) {
        std::fill(blk.begin(), blk.end(), 0);
        uint32_t c = 0;
        for (uint64_t j = 0; j < maxM0; j++) {
            int32_t v = g->adj0[(size_t)i * maxM0 + j];
            if (v < 0) break;
            std::memcpy(blk.data() + 4 + 4 * j, &v, 4);
            c++;
        }
        std::memcpy(blk.data(), &c, 4);
        std::memcpy(blk.data() + data_off, &g->vectors[(size_t)i * g->dim], 4 * (size_t)g->dim);
        uint64_t lab = g->labels.empty() ? (uint64_t)i : (uint64_t)g->labels[(size_t)i];
        std::memcpy(blk.data() + label_off, &lab, 8);
        f.write(blk.data(), (std::streamsize)size_per);
    }
    const uint64_t per_level = 4 + 4 * maxM;
    for (int64_t i = 0; i < g->n; i++) {
        int32_t L = g->levels.empty() ? 0 : g->levels[(size_t)i];
        uint32_t sz = (uint32_t)(L * per_level);
        wr(f, sz);
        for (int32_t l = 0; l < L; l++) {
            std::vector<char> row((size_t)per_level, 0);
            uint32_t c = 0;
            for (uint64_t j = 0; j < maxM; j++) {
                int32_t v = g->adj_up[(size_t)(g->upper_off[(size_t)i] + l) * maxM + j];
                if (v < 0) break;
                std::memcpy(row.data() + 4 + 4 * j, &v, 4);
                c++;
            }
            std::memcpy(row.data(), &c, 4);
            f.write(row.data(), (std::streamsize)per_level);
        }
    }
    return f ? GVS_OK : GVS_ERR_IO;
}

extern "C" int gvs_nsg_load(const char *path, const float *x, int64_t n, int32_t dim, int32_t metric,
                            gvs_host_graph **out)
{
    if (!path || !x || n <= 0 || dim <= 0 || !out) {
        gvs::set_error("gvs_nsg_load: invalid argument");
        return GVS_ERR_ARG;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        gvs::set_error("gvs_nsg_load: cannot open %s", path);
        return GVS_ERR_IO;
    }
    uint32_t width, ep;
    if (!rd(f, width) || !rd(f, ep) || width == 0 || width > 4096 || ep >= (uint64_t)n) {
        gvs::set_error("gvs_nsg_load: bad header");
        return GVS_ERR_IO;
    }
    gvs_host_graph *g = new gvs_host_graph();
    g->n = n;
    g->dim = dim;
    g->metric = metric;
    g->vectors.assign(x, x + (size_t)n * dim);
    g->r0 = (int32_t)width;
    g->adj0.assign((size_t)n * width, -1);
    for (int64_t i = 0; i < n; i++) {
        uint32_t k;
        if (!rd(f, k) || k > width) {
            delete g;
            gvs::set_error("gvs_nsg_load: bad row %lld", (long long)i);
            return GVS_ERR_IO;
        }
        for (uint32_t j = 0; j < k; j++) {
            uint32_t v;
            if (!rd(f, v) || v >= (uint64_t)n) {
                delete g;
                gvs::set_error("gvs_nsg_load: bad id in row %lld", (long long)i);
                return GVS_ERR_IO;
            }
            g->adj0[(size_t)i * width + j] = (int32_t)v;
        }
    }
    g->max_level = 0;
    g->entry = (int32_t)ep;
    g->levels.assign((size_t)n, 0);
    g->labels.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = i;
    *out = g;
    return GVS_OK;
}

extern "C" int gvs_host_graph_view(const gvs_host_graph *g, gvs_graph *v)
{
    if (!g || !v) {
        gvs::set_error("gvs_host_graph_view: invalid argument");
        return GVS_ERR_ARG;
    }
    std::memset(v, 0, sizeof *v);
    v->n = g->n;
    v->dim = g->dim;
    v->metric = g->metric;
    v->vectors = g->vectors.data();
    v->ld_vectors = g->dim;
    v->r0 = g->r0;
    v->adj0 = g->adj0.data();
    v->max_level = g->max_level;
    v->entry = g->entry;
    v->ru = g->ru;
    v->up_rows = g->ru > 0 ? (int64_t)(g->adj_up.size() / (size_t)g->ru) : 0;
    v->upper_off = g->upper_off.empty() ? nullptr : g->upper_off.data();
    v->adj_up = g->adj_up.empty() ? nullptr : g->adj_up.data();
    v->labels = g->labels.empty() ? nullptr : g->labels.data();
    return GVS_OK;
}

extern "C" void gvs_host_graph_free(gvs_host_graph *g) { delete g; }

extern "C" const char *gvs_version(void) { return "gvs-mi355x 0.1"; }
