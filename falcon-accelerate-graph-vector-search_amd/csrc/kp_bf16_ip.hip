// Search kernel instantiations with the initial-pool start (the NSG search):
// bf16 storage, inner product metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_pool_bf16_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_BF16, METRIC_IP, true>(rows);
    case 2: return pick_rows<2, STORE_BF16, METRIC_IP, true>(rows);
    case 3: return pick_rows<3, STORE_BF16, METRIC_IP, true>(rows);
    case 4: return pick_rows<4, STORE_BF16, METRIC_IP, true>(rows);
    case 6: return pick_rows<6, STORE_BF16, METRIC_IP, true>(rows);
    case 8: return pick_rows<8, STORE_BF16, METRIC_IP, true>(rows);
    default: return nullptr;
    }
}
