// Search kernel instantiations: bf16 storage, IP metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_bf16_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_BF16, METRIC_IP>(rows);
    case 2: return pick_rows<2, STORE_BF16, METRIC_IP>(rows);
    case 3: return pick_rows<3, STORE_BF16, METRIC_IP>(rows);
    case 4: return pick_rows<4, STORE_BF16, METRIC_IP>(rows);
    case 6: return pick_rows<6, STORE_BF16, METRIC_IP>(rows);
    case 8: return pick_rows<8, STORE_BF16, METRIC_IP>(rows);
    default: return nullptr;
    }
}

dfn_t gvs_descend_bf16_ip(int nch)
{
    switch (nch) {
    case 1: return k_descend<1, STORE_BF16, METRIC_IP>;
    case 2: return k_descend<2, STORE_BF16, METRIC_IP>;
    case 3: return k_descend<3, STORE_BF16, METRIC_IP>;
    case 4: return k_descend<4, STORE_BF16, METRIC_IP>;
    case 6: return k_descend<6, STORE_BF16, METRIC_IP>;
    case 8: return k_descend<8, STORE_BF16, METRIC_IP>;
    default: return nullptr;
    }
}
