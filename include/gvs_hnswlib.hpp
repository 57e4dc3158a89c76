/*
 * gvs_hnswlib.hpp - header-only C++ host API over the C ABI in gvs.h, with
 * the search surface of hnswlib's HierarchicalNSW so that code written
 * against hnswlib switches by changing the type it instantiates.
 *
 * Mirrors (paths relative to the reference root,
 * HNSWLIB = networked_FPGA/CPU_programs_oldest/hnswlib):
 *   HierarchicalNSW(space, location)        HNSWLIB/hnswalg.h:20-25, loadIndex :618-700
 *   setEf(ef)                               HNSWLIB/hnswalg.h:509-511
 *   searchKnn(query, k) -> max-heap, farthest on top
 *                                           HNSWLIB/hnswalg.h:1114-1168, hnswlib.h:158
 *   searchKnnCloserFirst(query, k)          HNSWLIB/hnswlib.h:160-185
 *   saveIndex(location)                     HNSWLIB/hnswalg.h:584-616
 * Errors are thrown as std::runtime_error with the library's message, as
 * hnswlib throws std::runtime_error ("Cannot open file", ...).
 *
 * Additions with no hnswlib counterpart: searchKnnBatch (one device launch
 * for a whole batch, the FPGA kernel's query_num / out_id / out_dist form,
 * INTER/vadd.cpp:11-95) and setDst (mc, mg of the DST scheduler,
 * INTER/scheduler.hpp:100-157).
 */
#ifndef GVS_HNSWLIB_HPP
#define GVS_HNSWLIB_HPP

#include <cstring>
#include <queue>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gvs.h"

namespace gvs_hnswlib {

typedef size_t labeltype;

inline void check(int rc)
{
    if (rc != GVS_OK) throw std::runtime_error(gvs_last_error());
}

class HierarchicalNSW {
public:
    /* Load an hnswlib .bin index (vectors of `dim` floats, metric
     * GVS_METRIC_L2 or GVS_METRIC_IP) and make it resident on the current
     * HIP device.  storage_mask: extra vector copies (1 << GVS_STORE_BF16,
     * 1 << GVS_STORE_FP8). */
    HierarchicalNSW(const std::string &location, int dim, int metric = GVS_METRIC_L2, uint32_t storage_mask = 0)
        : dim_(dim)
    {
        check(gvs_hnswlib_load(location.c_str(), dim, metric, &host_));
        gvs_graph view;
        int rc = gvs_host_graph_view(host_, &view);
        if (rc == GVS_OK) rc = gvs_index_create(&view, storage_mask | (1u << GVS_STORE_F32), &index_);
        if (rc != GVS_OK) {
            gvs_host_graph_free(host_);
            throw std::runtime_error(gvs_last_error());
        }
        std::memset(&params_, 0, sizeof params_);
        params_.ef = 10;  // hnswlib's default ef_
        params_.mc = 1;
        params_.mg = 1;
        params_.waves_per_query = 1;
        params_.storage = GVS_STORE_F32;
    }
    ~HierarchicalNSW()
    {
        gvs_index_destroy(index_);
        gvs_host_graph_free(host_);
    }
    HierarchicalNSW(const HierarchicalNSW &) = delete;
    HierarchicalNSW &operator=(const HierarchicalNSW &) = delete;

    void setEf(size_t ef) { params_.ef = (int32_t)ef; }
    size_t getEf() const { return (size_t)params_.ef; }
    /* delayed-synchronisation traversal: mc candidates per step, mg steps in flight */
    void setDst(int mc, int mg)
    {
        params_.mc = mc;
        params_.mg = mg;
    }
    /* traverse a low-precision copy (built via storage_mask), optionally
     * reranking the final ef candidates in exact fp32 */
    void setStorage(int storage, bool rerank)
    {
        params_.storage = storage;
        params_.rerank = rerank ? 1 : 0;
    }

    std::priority_queue<std::pair<float, labeltype>> searchKnn(const void *query_data, size_t k) const
    {
        std::vector<int64_t> lab(k);
        std::vector<float> dist(k);
        searchKnnBatch((const float *)query_data, 1, k, lab.data(), dist.data());
        std::priority_queue<std::pair<float, labeltype>> result;
        for (size_t i = 0; i < k; i++)
            if (lab[i] >= 0) result.emplace(dist[i], (labeltype)lab[i]);
        return result;
    }

    std::vector<std::pair<float, labeltype>> searchKnnCloserFirst(const void *query_data, size_t k) const
    {
        auto q = searchKnn(query_data, k);
        std::vector<std::pair<float, labeltype>> out(q.size());
        for (size_t i = q.size(); i-- > 0;) {
            out[i] = q.top();
            q.pop();
        }
        return out;
    }

    /* nq host queries (row stride dim) -> labels[nq*k], dists[nq*k] closer
     * first; missing results are label -1 / dist +inf. */
    void searchKnnBatch(const float *queries, size_t nq, size_t k, int64_t *labels, float *dists,
                        int32_t *stats = nullptr) const
    {
        gvs_params p = params_;
        p.k = (int32_t)k;
        check(gvs_search_host(index_, queries, (int64_t)nq, dim_, nullptr, &p, labels, dists, stats));
    }

    void saveIndex(const std::string &location) const { check(gvs_hnswlib_save(host_, location.c_str())); }

    gvs_index *index() const { return index_; }

private:
    int dim_;
    gvs_host_graph *host_ = nullptr;
    gvs_index *index_ = nullptr;
    gvs_params params_;
};

}  // namespace gvs_hnswlib

#endif
