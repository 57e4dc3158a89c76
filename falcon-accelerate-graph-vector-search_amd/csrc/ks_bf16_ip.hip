// Serving kernel instantiations: bf16 storage (fp32 accumulation, optional fp32 rerank), IP metric.
#include "serve_kernel.hpp"

using namespace gvsk;

sfn_t gvs_serve_kernel_bf16_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_serve_rows<1, STORE_BF16, METRIC_IP>(rows);
    case 2: return pick_serve_rows<2, STORE_BF16, METRIC_IP>(rows);
    case 3: return pick_serve_rows<3, STORE_BF16, METRIC_IP>(rows);
    case 4: return pick_serve_rows<4, STORE_BF16, METRIC_IP>(rows);
    case 6: return pick_serve_rows<6, STORE_BF16, METRIC_IP>(rows);
    case 8: return pick_serve_rows<8, STORE_BF16, METRIC_IP>(rows);
    default: return nullptr;
    }
}
