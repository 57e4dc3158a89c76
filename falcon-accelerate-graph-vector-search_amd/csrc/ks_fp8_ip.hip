// Serving kernel instantiations: fp8 storage (fp32 accumulation, optional fp32 rerank), IP metric.
#include "serve_kernel.hpp"

using namespace gvsk;

sfn_t gvs_serve_kernel_fp8_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_serve_rows<1, STORE_FP8, METRIC_IP>(rows);
    case 2: return pick_serve_rows<2, STORE_FP8, METRIC_IP>(rows);
    case 3: return pick_serve_rows<3, STORE_FP8, METRIC_IP>(rows);
    case 4: return pick_serve_rows<4, STORE_FP8, METRIC_IP>(rows);
    default: return nullptr;
    }
}
