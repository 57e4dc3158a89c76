// Search kernel instantiations with the initial-pool start (the NSG search):
// fp8 storage, inner product metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_pool_fp8_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_FP8, METRIC_IP, true>(rows);
    case 2: return pick_rows<2, STORE_FP8, METRIC_IP, true>(rows);
    case 3: return pick_rows<3, STORE_FP8, METRIC_IP, true>(rows);
    case 4: return pick_rows<4, STORE_FP8, METRIC_IP, true>(rows);
    default: return nullptr;
    }
}
