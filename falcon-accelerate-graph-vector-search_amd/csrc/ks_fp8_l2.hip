// Serving kernel instantiations: fp8 storage (fp32 accumulation, optional fp32 rerank), L2 metric.
#include "serve_kernel.hpp"

using namespace gvsk;

sfn_t gvs_serve_kernel_fp8_l2(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_serve_rows<1, STORE_FP8, METRIC_L2>(rows);
    case 2: return pick_serve_rows<2, STORE_FP8, METRIC_L2>(rows);
    case 3: return pick_serve_rows<3, STORE_FP8, METRIC_L2>(rows);
    case 4: return pick_serve_rows<4, STORE_FP8, METRIC_L2>(rows);
    default: return nullptr;
    }
}
