// Search kernel instantiations: fp8 storage, L2 metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_fp8_l2(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_FP8, METRIC_L2>(rows);
    case 2: return pick_rows<2, STORE_FP8, METRIC_L2>(rows);
    case 3: return pick_rows<3, STORE_FP8, METRIC_L2>(rows);
    case 4: return pick_rows<4, STORE_FP8, METRIC_L2>(rows);
    default: return nullptr;
    }
}

dfn_t gvs_descend_fp8_l2(int nch)
{
    switch (nch) {
    case 1: return k_descend<1, STORE_FP8, METRIC_L2>;
    case 2: return k_descend<2, STORE_FP8, METRIC_L2>;
    case 3: return k_descend<3, STORE_FP8, METRIC_L2>;
    case 4: return k_descend<4, STORE_FP8, METRIC_L2>;
    default: return nullptr;
    }
}
