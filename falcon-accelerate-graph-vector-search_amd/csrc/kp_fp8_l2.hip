// Search kernel instantiations with the initial-pool start (the NSG search):
// fp8 storage, L2 metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_pool_fp8_l2(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_FP8, METRIC_L2, true>(rows);
    case 2: return pick_rows<2, STORE_FP8, METRIC_L2, true>(rows);
    case 3: return pick_rows<3, STORE_FP8, METRIC_L2, true>(rows);
    case 4: return pick_rows<4, STORE_FP8, METRIC_L2, true>(rows);
    default: return nullptr;
    }
}
