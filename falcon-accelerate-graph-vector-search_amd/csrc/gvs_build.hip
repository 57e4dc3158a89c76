// GPU HNSW construction with hnswlib's insertion rules
// (networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:982-1112 addPoint),
// the index build the reference runs on the CPU
// (vector_search_baselines/scripts_hnsw/construct_and_search_hnsw.py:126-148).
//
// Nodes are inserted in index order, in batches [a, b): every node of a
// batch is inserted against the graph as it stood at the batch start, and
// the batches grow with the graph (b - a = ratio * a, at most max_batch), so
// a node misses as links only the few nodes of its own batch.  Per batch:
//   host    the levels >= 1 of the batch's upper nodes (about 1/M of them):
//           descent, construction search and heuristic per level, reverse
//           links grouped by target (hnsw_build.cpp, UpperBatch); their
//           level-1 result is their layer-0 entry;
//   device  layer 0 for the whole batch: the search kernel with
//           ef = k = efConstruction from that entry (the others descend the
//           upper levels on the device), k_build_select (the heuristic, M
//           links, the node's row) and, after a radix sort of the
//           reverse-link keys by target, k_build_link (append or re-prune).
// A batch of one node is addPoint; with hnswlib's scalar summation order
// (exact_order) the graph is then hnswlib's single-thread graph bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "build_kernels.hpp"
#include "gvs_internal.hpp"

using namespace gvsk;

namespace gvs {
struct UpperBuild;
UpperBuild *upper_build_create(const float *x, int64_t n, int dim, int metric, int M, int efc, uint64_t seed,
                               bool exact);
void upper_build_free(UpperBuild *ub);
bool upper_build_batch(UpperBuild *ub, int64_t a, int64_t b, int nthreads, std::vector<int32_t> &l0_entry);
int upper_build_entry(const UpperBuild *ub);
int upper_build_maxlevel(const UpperBuild *ub);
int upper_build_final_maxlevel(const UpperBuild *ub);
int64_t upper_build_final_entry(const UpperBuild *ub);
const std::vector<int64_t> &upper_build_offsets(const UpperBuild *ub);
const std::vector<int32_t> &upper_build_rows(const UpperBuild *ub);
int64_t upper_build_nrows(const UpperBuild *ub);
const std::vector<int> &upper_build_levels(const UpperBuild *ub);
int upper_build_efc(const UpperBuild *ub);
double upper_build_mult(const UpperBuild *ub);
int search_rows(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq, const int32_t *entries,
                const gvs_params *p, int64_t *labels, float *dists, hipStream_t stream);
}  // namespace gvs

namespace {

typedef void (*bfn_t)(KParams, BuildArgs);

template <int STORE, int METRIC>
bfn_t pick_build(int nch, bool link)
{
#define GVS_BUILD_CASE(N) \
    case N: return link ? k_build_link<N, STORE, METRIC> : k_build_select<N, STORE, METRIC>;
    switch (nch) {
        GVS_BUILD_CASE(1)
        GVS_BUILD_CASE(2)
        GVS_BUILD_CASE(3)
        GVS_BUILD_CASE(4)
        GVS_BUILD_CASE(6)
        GVS_BUILD_CASE(8)
        GVS_BUILD_CASE(12)
        GVS_BUILD_CASE(16)
    default: return nullptr;
    }
#undef GVS_BUILD_CASE
}

bfn_t pick(int nch, bool seq, bool l2, bool link)
{
    if (seq) return l2 ? pick_build<STORE_F32_SEQ, METRIC_L2>(nch, link) : pick_build<STORE_F32_SEQ, METRIC_IP>(nch, link);
    return l2 ? pick_build<STORE_F32, METRIC_L2>(nch, link) : pick_build<STORE_F32, METRIC_IP>(nch, link);
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
    bool alloc(size_t bytes, const char *what) { return gvs::hip_ok(hipMalloc(&p, std::max<size_t>(bytes, 16)), what); }
};

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

extern "C" int gvs_hnsw_build_device(const float *x, int64_t n, int32_t dim, int32_t metric, int32_t M,
                                     int32_t ef_construction, uint64_t seed, const gvs_build_config *cfg,
                                     gvs_host_graph **out, double stats[8])
{
    if (!x || n <= 0 || n > INT32_MAX || dim <= 0 || dim > 512 || M < 2 || M > WAVE / 2 || ef_construction < 1 ||
        !out || (metric != GVS_METRIC_L2 && metric != GVS_METRIC_IP)) {
        gvs::set_error("gvs_hnsw_build_device: invalid argument (2 <= M <= 32, dim <= 512)");
        return GVS_ERR_ARG;
    }
    const int efc = std::max(ef_construction, M);
    if (efc > BUILD_CAP) {
        gvs::set_error("gvs_hnsw_build_device: ef_construction must be <= %d", BUILD_CAP);
        return GVS_ERR_ARG;
    }
    gvs_build_config c{};
    if (cfg) c = *cfg;
    const bool exact = c.exact_order != 0;
    const double ratio = c.batch_ratio > 0.0 ? c.batch_ratio : (c.batch_ratio < 0.0 ? 0.0 : 0.01);
    const int64_t max_batch = c.max_batch > 0 ? c.max_batch : 32768;
    const int threads = c.host_threads > 0 ? c.host_threads : 16;
    const Clock::time_point t_start = Clock::now();
    double t_upper = 0.0, t_search = 0.0, t_link = 0.0;
    int64_t nbatches = 0;

    gvs::UpperBuild *ub = gvs::upper_build_create(x, n, dim, metric, M, efc, seed, exact);
    struct UbFree {
        gvs::UpperBuild *u;
        ~UbFree() { gvs::upper_build_free(u); }
    } ub_free{ub};
    const int r0 = 2 * M;

    // device index over the vectors, with empty rows at every level
    gvs_index *ix = nullptr;
    {
        std::vector<int32_t> adj0((size_t)n * r0, -1);
        gvs_graph g;
        std::memset(&g, 0, sizeof g);
        g.n = n;
        g.dim = dim;
        g.metric = metric;
        g.vectors = x;
        g.ld_vectors = dim;
        g.r0 = r0;
        g.adj0 = adj0.data();
        g.max_level = gvs::upper_build_final_maxlevel(ub);
        g.entry = (int32_t)gvs::upper_build_final_entry(ub);
        g.ru = M;
        g.up_rows = gvs::upper_build_nrows(ub);
        g.upper_off = g.max_level > 0 ? gvs::upper_build_offsets(ub).data() : nullptr;
        g.adj_up = g.max_level > 0 ? gvs::upper_build_rows(ub).data() : nullptr;
        const int rc = gvs_index_create(&g, 1u << GVS_STORE_F32, &ix);
        if (rc) return rc;
    }
    struct IxFree {
        gvs_index *p;
        ~IxFree() { (void)gvs_index_destroy(p); }
    } ix_free{ix};
    const int nch = ix->dim_pad / 32;
    const bool l2 = metric == GVS_METRIC_L2;
    const bfn_t f_sel = pick(nch, exact, l2, false), f_link = pick(nch, exact, l2, true);
    if (!f_sel || !f_link) {
        gvs::set_error("gvs_hnsw_build_device: dim %d (padded %d) not compiled", dim, ix->dim_pad);
        return GVS_ERR_UNSUPPORTED;
    }

    DevBuf d_cand, d_cdist, d_ent, d_pairs, d_sorted, d_tmp;
    const int64_t cap = std::min<int64_t>(max_batch, n);
    if (!d_cand.alloc(sizeof(int64_t) * cap * efc, "hipMalloc candidates") ||
        !d_cdist.alloc(sizeof(float) * cap * efc, "hipMalloc candidate distances") ||
        !d_ent.alloc(sizeof(int32_t) * cap, "hipMalloc entries") ||
        !d_pairs.alloc(sizeof(uint64_t) * cap * M, "hipMalloc link keys") ||
        !d_sorted.alloc(sizeof(uint64_t) * cap * M, "hipMalloc sorted link keys"))
        return GVS_ERR_NOMEM;
    int key_bits = 32;
    while (key_bits < 64 && (((int64_t)1) << (key_bits - 32)) < n) key_bits++;
    size_t tmp_bytes = 0;
    if (!gvs::hip_ok(rocprim::radix_sort_keys(nullptr, tmp_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                              (size_t)(cap * M), 0, key_bits, nullptr),
                     "radix sort size") ||
        !d_tmp.alloc(tmp_bytes, "hipMalloc sort storage"))
        return GVS_ERR_NOMEM;

    KParams P;
    std::memset(&P, 0, sizeof P);
    P.vec = ix->d_f32;
    P.vec_f32 = ix->d_f32;
    P.row_elems = ix->dim_pad;
    P.row_elems_f32 = ix->dim_pad;
    P.dim = dim;
    P.r0 = r0;
    P.n = n;

    gvs_params sp;
    std::memset(&sp, 0, sizeof sp);
    sp.k = efc;
    sp.ef = efc;
    sp.mc = 1;
    sp.mg = 1;
    sp.waves_per_query = 1;
    sp.storage = GVS_STORE_F32;
    sp.flags = exact ? GVS_FLAG_HNSWLIB_ORDER : 0;

    std::vector<int32_t> l0_entry;
    const int64_t up_words = gvs::upper_build_nrows(ub) * M;
    for (int64_t a = 0; a < n;) {
        const int64_t size = a == 0 ? 1 : std::max<int64_t>(1, std::min<int64_t>(max_batch, (int64_t)(a * ratio)));
        const int64_t b = std::min(n, a + size);
        const Clock::time_point t0 = Clock::now();
        const bool changed = gvs::upper_build_batch(ub, a, b, threads, l0_entry);
        if (changed && up_words > 0 &&
            !gvs::hip_ok(hipMemcpy(ix->d_adj_up, gvs::upper_build_rows(ub).data(), sizeof(int32_t) * up_words,
                                   hipMemcpyHostToDevice),
                         "copy upper rows"))
            return GVS_ERR_HIP;
        ix->max_level = gvs::upper_build_maxlevel(ub);
        ix->entry = gvs::upper_build_entry(ub);
        const Clock::time_point t1 = Clock::now();
        t_upper += secs(t0, t1);
        // layer 0 (the first node is the entry point and has no links yet)
        const int64_t a0 = a == 0 ? 1 : a;
        const int64_t nb = b - a0;
        if (nb > 0) {
            if (!gvs::hip_ok(hipMemcpy(d_ent.p, l0_entry.data() + (a0 - a), sizeof(int32_t) * nb,
                                       hipMemcpyHostToDevice),
                             "copy entries"))
                return GVS_ERR_HIP;
            int rc = gvs::search_rows(ix, ix->d_f32 + a0 * ix->dim_pad, nb, ix->dim_pad, (const int32_t *)d_ent.p,
                                      &sp, (int64_t *)d_cand.p, (float *)d_cdist.p, nullptr);
            if (rc) return rc;
            const Clock::time_point t2 = Clock::now();
            t_search += secs(t1, t2);
            BuildArgs A;
            A.first = a0;
            A.nb = (int)nb;
            A.efc = efc;
            A.cand = (const int64_t *)d_cand.p;
            A.cdist = (const float *)d_cdist.p;
            A.M = M;
            A.adj0 = ix->d_adj0;
            A.pairs = (uint64_t *)d_pairs.p;
            A.npairs = nb * M;
            hipLaunchKernelGGL(f_sel, dim3((unsigned)nb), dim3(WAVE), 0, nullptr, P, A);
            if (!gvs::hip_ok(hipGetLastError(), "launch k_build_select")) return GVS_ERR_HIP;
            size_t tb = tmp_bytes;
            if (!gvs::hip_ok(rocprim::radix_sort_keys(d_tmp.p, tb, (uint64_t *)d_pairs.p, (uint64_t *)d_sorted.p,
                                                      (size_t)(nb * M), 0, key_bits, nullptr),
                             "radix sort link keys"))
                return GVS_ERR_HIP;
            A.pairs = (uint64_t *)d_sorted.p;
            hipLaunchKernelGGL(f_link, dim3((unsigned)(nb * M)), dim3(WAVE), 0, nullptr, P, A);
            if (!gvs::hip_ok(hipGetLastError(), "launch k_build_link") ||
                !gvs::hip_ok(hipDeviceSynchronize(), "build batch"))
                return GVS_ERR_HIP;
            t_link += secs(t2, Clock::now());
        }
        nbatches++;
        a = b;
    }

    // the host graph
    gvs_host_graph *g = new gvs_host_graph();
    g->n = n;
    g->dim = dim;
    g->metric = metric;
    g->vectors.assign(x, x + (size_t)n * dim);
    g->r0 = r0;
    g->adj0.resize((size_t)n * r0);
    if (!gvs::hip_ok(hipMemcpy(g->adj0.data(), ix->d_adj0, sizeof(int32_t) * n * r0, hipMemcpyDeviceToHost),
                     "copy rows")) {
        delete g;
        return GVS_ERR_HIP;
    }
    g->max_level = gvs::upper_build_maxlevel(ub);
    g->entry = gvs::upper_build_entry(ub);
    g->ru = M;
    g->levels.assign(gvs::upper_build_levels(ub).begin(), gvs::upper_build_levels(ub).end());
    g->upper_off = gvs::upper_build_offsets(ub);
    g->adj_up = gvs::upper_build_rows(ub);
    g->labels.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = i;
    g->M = M;
    g->ef_construction = gvs::upper_build_efc(ub);
    g->mult = gvs::upper_build_mult(ub);
    *out = g;
    if (stats) {
        stats[0] = secs(t_start, Clock::now());
        stats[1] = t_upper;
        stats[2] = t_search;
        stats[3] = t_link;
        stats[4] = (double)nbatches;
        stats[5] = stats[6] = stats[7] = 0.0;
    }
    return GVS_OK;
}
