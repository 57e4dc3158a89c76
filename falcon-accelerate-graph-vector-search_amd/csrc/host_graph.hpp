// Host-side graph object shared by the builder, the format readers and the
// device index.  Flat layout described in include/gvs.h (gvs_graph).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "gvs.h"

struct gvs_host_graph {
    int64_t n = 0;
    int32_t dim = 0;
    int32_t metric = GVS_METRIC_L2;
    std::vector<float> vectors;      // n * dim
    int32_t r0 = 0;
    std::vector<int32_t> adj0;       // n * r0, -1 padded
    int32_t max_level = 0;
    int32_t entry = 0;
    int32_t ru = 0;
    std::vector<int64_t> upper_off;  // n (or empty when max_level == 0)
    std::vector<int32_t> adj_up;     // rows * ru
    std::vector<int32_t> levels;     // n
    std::vector<int64_t> labels;     // n
    // hnswlib header fields kept for saveIndex
    int32_t M = 0;
    int32_t ef_construction = 0;
    double mult = 0.0;
};

namespace gvs {
void set_error(const char *fmt, ...);
}
