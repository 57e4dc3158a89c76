// On-disk formats and host-graph plumbing.
//   hnswlib .bin  : networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:584-700
//                   (96-byte header, level-0 blocks of count + maxM0 ids +
//                   vector + size_t label, then per node a u32 byte count and
//                   its upper-level lists); parsed the same way by
//                   vector_search_baselines/scripts_hnsw/hnsw.py:139-262.
//   NSG .nsg      : vector_search_baselines/scripts_nsg/nsg_to_FPGA.py:31-47
//                   (u32 width, u32 ep, then per node u32 k and k u32 ids).
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
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

// ---- FPGA-layout index directory ---------------------------------------
// The reference's converted format (vector_search_baselines/scripts_hnsw/hnsw.py:264-377,
// scripts_nsg/nsg_to_FPGA.py:56-137), all little-endian:
//   meta.bin                HNSW: n, maxlevel, entry, maxM, maxM0 (uint32 each)
//                           NSG:  n, entry, width
//   ground_links_<nc>_chan_<c>.bin    node i in channel i % nc, record i / nc:
//                           [count u32 + 60 B pad][r0 x u32 ids, 0 when unused], 64 B aligned
//   ground_vectors_<nc>_chan_<c>.bin  [dim x f32, 64 B aligned][visited i32 = -1 + 60 B pad]
//   ground_labels.bin       n x u32 (HNSW; absent = identity)
//   upper_links.bin         per node with level L >= 1, levels 1..L:
//                           [count u32 + 60 B pad][maxM x u32], 64 B aligned
//   upper_links_pointers.bin n x u64 byte offset of the node's first upper record
namespace {
inline int64_t rec64(int64_t ints) { return (ints + 15) / 16 * 64 + 64; }

bool read_file(const std::string &path, std::vector<unsigned char> &buf)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    const std::streamoff sz = f.tellg();
    f.seekg(0, std::ios::beg);
    buf.resize((size_t)sz);
    return sz == 0 || (bool)f.read((char *)buf.data(), sz);
}

uint32_t rd32(const unsigned char *p)
{
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

std::string chan_name(const char *dir, const char *what, int nc, int c)
{
    return std::string(dir) + "/" + what + "_" + std::to_string(nc) + "_chan_" + std::to_string(c) + ".bin";
}
}  // namespace

extern "C" int gvs_fpga_load(const char *dir, int32_t kind, int32_t dim, int32_t nchannels, int32_t metric,
                             gvs_host_graph **out)
{
    if (!dir || !out || (kind != 0 && kind != 1) || dim <= 0 || nchannels < 1 ||
        (metric != GVS_METRIC_L2 && metric != GVS_METRIC_IP)) {
        gvs::set_error("gvs_fpga_load: invalid argument");
        return GVS_ERR_ARG;
    }
    const std::string d(dir);
    std::vector<unsigned char> meta;
    if (!read_file(d + "/meta.bin", meta) || meta.size() < (kind == 0 ? 20u : 12u)) {
        gvs::set_error("gvs_fpga_load: cannot read %s/meta.bin", dir);
        return GVS_ERR_IO;
    }
    const int64_t n = rd32(&meta[0]);
    int32_t maxlevel = 0, entry, maxM = 0, r0;
    if (kind == 0) {
        maxlevel = (int32_t)rd32(&meta[4]);
        entry = (int32_t)rd32(&meta[8]);
        maxM = (int32_t)rd32(&meta[12]);
        r0 = (int32_t)rd32(&meta[16]);
    } else {
        entry = (int32_t)rd32(&meta[4]);
        r0 = (int32_t)rd32(&meta[8]);
    }
    if (n <= 0 || n > INT32_MAX || entry < 0 || entry >= n || r0 <= 0 || r0 > 4096 || maxlevel < 0 ||
        maxlevel > 64 || (maxlevel > 0 && (maxM <= 0 || maxM > 4096))) {
        gvs::set_error("gvs_fpga_load: %s/meta.bin has out-of-range fields", dir);
        return GVS_ERR_IO;
    }
    gvs_host_graph *g = new gvs_host_graph();
    g->n = n;
    g->dim = dim;
    g->metric = metric;
    g->r0 = r0;
    g->entry = entry;
    g->vectors.assign((size_t)n * dim, 0.0f);
    g->adj0.assign((size_t)n * r0, -1);
    g->levels.assign((size_t)n, 0);
    g->labels.resize((size_t)n);
    const int64_t lrec = rec64(r0), vrec = rec64(dim);
    const int nc = nchannels;
    for (int c = 0; c < nc; c++) {
        const int64_t cnt = n / nc + (c < n % nc ? 1 : 0);
        std::vector<unsigned char> lb, vb;
        if (!read_file(chan_name(dir, "ground_links", nc, c), lb) || (int64_t)lb.size() != cnt * lrec ||
            !read_file(chan_name(dir, "ground_vectors", nc, c), vb) || (int64_t)vb.size() != cnt * vrec) {
            delete g;
            gvs::set_error("gvs_fpga_load: channel %d of %d: missing file or wrong size (%s)", c, nc,
                           chan_name(dir, "ground_links", nc, c).c_str());
            return GVS_ERR_IO;
        }
        for (int64_t j = 0; j < cnt; j++) {
            const int64_t i = j * nc + c;
            const unsigned char *lr = &lb[(size_t)(j * lrec)];
            const uint32_t links = rd32(lr);
            if (links > (uint32_t)r0) {
                delete g;
                gvs::set_error("gvs_fpga_load: node %lld has %u links > %d", (long long)i, links, r0);
                return GVS_ERR_IO;
            }
            for (uint32_t t = 0; t < links; t++) {
                const uint32_t u = rd32(lr + 64 + 4 * t);
                if (u >= (uint64_t)n) {
                    delete g;
                    gvs::set_error("gvs_fpga_load: node %lld links to %u >= n", (long long)i, u);
                    return GVS_ERR_IO;
                }
                g->adj0[(size_t)(i * r0 + t)] = (int32_t)u;
            }
            std::memcpy(&g->vectors[(size_t)(i * dim)], &vb[(size_t)(j * vrec)], sizeof(float) * (size_t)dim);
        }
    }
    std::vector<unsigned char> lab;
    if (kind == 0 && read_file(d + "/ground_labels.bin", lab) && (int64_t)lab.size() == 4 * n) {
        for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = rd32(&lab[(size_t)(4 * i)]);
    } else {
        for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = i;
    }
    g->max_level = maxlevel;
    if (kind == 0 && maxlevel > 0) {
        std::vector<unsigned char> ub, pb;
        if (!read_file(d + "/upper_links.bin", ub) || !read_file(d + "/upper_links_pointers.bin", pb) ||
            (int64_t)pb.size() != 8 * n) {
            delete g;
            gvs::set_error("gvs_fpga_load: cannot read the upper links in %s", dir);
            return GVS_ERR_IO;
        }
        const int64_t urec = rec64(maxM);
        g->ru = maxM;
        g->upper_off.assign((size_t)n, -1);
        std::vector<uint64_t> ptr((size_t)n + 1);
        std::memcpy(ptr.data(), pb.data(), 8 * (size_t)n);
        ptr[(size_t)n] = ub.size();
        int64_t rows = 0;
        for (int64_t i = 0; i < n; i++) {
            if (ptr[(size_t)i + 1] < ptr[(size_t)i] || (ptr[(size_t)i + 1] - ptr[(size_t)i]) % (uint64_t)urec) {
                delete g;
                gvs::set_error("gvs_fpga_load: upper_links_pointers.bin is not increasing by whole records");
                return GVS_ERR_IO;
            }
            const int L = (int)((ptr[(size_t)i + 1] - ptr[(size_t)i]) / (uint64_t)urec);
            g->levels[(size_t)i] = L;
            if (L > 0) {
                g->upper_off[(size_t)i] = rows;
                rows += L;
            }
        }
        g->adj_up.assign((size_t)(rows * maxM), -1);
        for (int64_t i = 0; i < n; i++) {
            for (int l = 0; l < g->levels[(size_t)i]; l++) {
                const unsigned char *r = &ub[(size_t)(ptr[(size_t)i] + (uint64_t)(l * urec))];
                const uint32_t links = std::min<uint32_t>(rd32(r), (uint32_t)maxM);
                for (uint32_t t = 0; t < links; t++) {
                    const uint32_t u = rd32(r + 64 + 4 * t);
                    if (u >= (uint64_t)n) {
                        delete g;
                        gvs::set_error("gvs_fpga_load: upper link to %u >= n", u);
                        return GVS_ERR_IO;
                    }
                    g->adj_up[(size_t)((g->upper_off[(size_t)i] + l) * maxM + t)] = (int32_t)u;
                }
            }
        }
    } else {
        g->max_level = 0;
    }
    g->M = maxM;
    *out = g;
    return GVS_OK;
}

extern "C" int gvs_fpga_save(const gvs_host_graph *g, const char *dir, int32_t nchannels)
{
    if (!g || !dir || nchannels < 1) {
        gvs::set_error("gvs_fpga_save: invalid argument");
        return GVS_ERR_ARG;
    }
    const bool hnsw = g->max_level > 0;
    const std::string d(dir);
    auto put32 = [](std::vector<unsigned char> &b, uint32_t v) {
        const size_t o = b.size();
        b.resize(o + 4);
        std::memcpy(&b[o], &v, 4);
    };
    auto pad64 = [](std::vector<unsigned char> &b) { b.resize((b.size() + 63) / 64 * 64, 0); };
    auto write = [&](const std::string &path, const std::vector<unsigned char> &b) {
        std::ofstream f(path, std::ios::binary);
        f.write((const char *)b.data(), (std::streamsize)b.size());
        return (bool)f;
    };
    std::vector<unsigned char> meta;
    put32(meta, (uint32_t)g->n);
    if (hnsw) {
        put32(meta, (uint32_t)g->max_level);
        put32(meta, (uint32_t)g->entry);
        put32(meta, (uint32_t)g->ru);
        put32(meta, (uint32_t)g->r0);
    } else {
        put32(meta, (uint32_t)g->entry);
        put32(meta, (uint32_t)g->r0);
    }
    bool ok = write(d + "/meta.bin", meta);
    const int nc = nchannels;
    std::vector<std::vector<unsigned char>> lb((size_t)nc), vb((size_t)nc);
    for (int64_t i = 0; i < g->n && ok; i++) {
        std::vector<unsigned char> &L = lb[(size_t)(i % nc)], &V = vb[(size_t)(i % nc)];
        const int32_t *row = &g->adj0[(size_t)(i * g->r0)];
        uint32_t cnt = 0;
        while ((int)cnt < g->r0 && row[cnt] >= 0) cnt++;
        put32(L, cnt);
        L.resize(L.size() + 60, 0);
        for (int t = 0; t < g->r0; t++) put32(L, t < (int)cnt ? (uint32_t)row[t] : 0u);
        pad64(L);
        const size_t o = V.size();
        V.resize(o + sizeof(float) * (size_t)g->dim);
        std::memcpy(&V[o], &g->vectors[(size_t)(i * g->dim)], sizeof(float) * (size_t)g->dim);
        pad64(V);
        put32(V, 0xFFFFFFFFu);  // the FPGA's visited flag, -1
        V.resize(V.size() + 60, 0);
    }
    for (int c = 0; c < nc && ok; c++)
        ok = write(chan_name(dir, "ground_links", nc, c), lb[(size_t)c]) &&
             write(chan_name(dir, "ground_vectors", nc, c), vb[(size_t)c]);
    if (ok && hnsw) {
        std::vector<unsigned char> lab, ub, pb;
        for (int64_t i = 0; i < g->n; i++)
            put32(lab, (uint32_t)(g->labels.empty() ? i : g->labels[(size_t)i]));
        for (int64_t i = 0; i < g->n; i++) {
            const uint64_t addr = ub.size();
            const size_t o = pb.size();
            pb.resize(o + 8);
            std::memcpy(&pb[o], &addr, 8);
            const int L = g->levels.empty() ? 0 : g->levels[(size_t)i];
            for (int l = 0; l < L; l++) {
                const int32_t *r = &g->adj_up[(size_t)((g->upper_off[(size_t)i] + l) * g->ru)];
                uint32_t cnt = 0;
                while ((int)cnt < g->ru && r[cnt] >= 0) cnt++;
                put32(ub, cnt);
                ub.resize(ub.size() + 60, 0);
                for (int t = 0; t < g->ru; t++) put32(ub, t < (int)cnt ? (uint32_t)r[t] : 0u);
                pad64(ub);
            }
        }
        ok = write(d + "/ground_labels.bin", lab) && write(d + "/upper_links.bin", ub) &&
             write(d + "/upper_links_pointers.bin", pb);
    }
    if (!ok) {
        gvs::set_error("gvs_fpga_save: cannot write into %s", dir);
        return GVS_ERR_IO;
    }
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
