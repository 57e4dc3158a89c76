// Search kernel instantiations with the initial-pool start (the NSG search):
// f32 storage, inner product metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_pool_f32_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_F32, METRIC_IP, true>(rows);
    case 2: return pick_rows<2, STORE_F32, METRIC_IP, true>(rows);
    case 3: return pick_rows<3, STORE_F32, METRIC_IP, true>(rows);
    case 4: return pick_rows<4, STORE_F32, METRIC_IP, true>(rows);
    case 6: return pick_rows<6, STORE_F32, METRIC_IP, true>(rows);
    case 8: return pick_rows<8, STORE_F32, METRIC_IP, true>(rows);
    case 12: return pick_rows<12, STORE_F32, METRIC_IP, true>(rows);
    case 16: return pick_rows<16, STORE_F32, METRIC_IP, true>(rows);
    default: return nullptr;
    }
}
