// Internal declarations shared by the device-side translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "host_graph.hpp"

struct gvs_index {
    int device = 0;
    int cu_count = 0;
    int64_t n = 0;
    int dim = 0, dim_pad = 0, metric = 0;
    int r0 = 0, ru = 0, max_level = 0, entry = 0;
    int64_t up_rows = 0;
    uint32_t storage_mask = 1;
    float *d_f32 = nullptr;
    uint16_t *d_bf16 = nullptr;
    uint8_t *d_fp8 = nullptr;
    float *d_scale = nullptr;
    int32_t *d_adj0 = nullptr;
    int32_t *d_adj_up = nullptr;
    int64_t *d_upper_off = nullptr;
    int64_t *d_labels = nullptr;
    // A fork (gvs_index_fork) shares the arrays above with its root index and
    // owns only what follows; the root counts its open forks.
    gvs_index *parent = nullptr;
    int forks = 0;
    // scratch: per-slot visited tables, work-queue counters, per-query status
    uint32_t *d_vtab = nullptr;
    size_t vtab_words = 0;
    uint32_t *d_counters = nullptr;
    int32_t *d_status = nullptr;
    int64_t status_cap = 0;
    double *d_xnorm64 = nullptr;   // |x|^2 in fp64 for the brute-force GT kernel
    // region order scratch (grown on demand): entries, keys, permutation, sort storage
    int32_t *d_rep = nullptr;
    uint64_t *d_rkey = nullptr;    // 2 * rcap
    int32_t *d_rperm = nullptr;    // 2 * rcap
    void *d_rtmp = nullptr;
    int64_t rcap = 0;
    size_t rtmp_bytes = 0;
};

namespace gvs {
bool hip_ok(hipError_t e, const char *what);

struct LaunchPlan {
    void *fn = nullptr;
    void *dfn = nullptr;           // upper-layer descent kernel (region order)
    int ef_cap = 0, lcap = 0, blog2 = 0, vlog2 = 0, hlog2 = 0, hnarrow = 0, hbits = 0, lds_per_wave = 0, wpb = 1;
    int mg = 1;                    // DST groups in flight (sizes the LDS ring of selected groups)
    int64_t resident_waves = 0;
};
}  // namespace gvs
