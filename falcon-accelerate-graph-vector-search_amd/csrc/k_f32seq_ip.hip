// Search kernel instantiations: fp32 rows in hnswlib's scalar summation order, IP metric.
#include "search_kernels.hpp"

using namespace gvsk;

kfn_t gvs_kernel_f32seq_ip(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_rows<1, STORE_F32_SEQ, METRIC_IP>(rows);
    case 2: return pick_rows<2, STORE_F32_SEQ, METRIC_IP>(rows);
    case 3: return pick_rows<3, STORE_F32_SEQ, METRIC_IP>(rows);
    case 4: return pick_rows<4, STORE_F32_SEQ, METRIC_IP>(rows);
    case 6: return pick_rows<6, STORE_F32_SEQ, METRIC_IP>(rows);
    case 8: return pick_rows<8, STORE_F32_SEQ, METRIC_IP>(rows);
    case 12: return pick_rows<12, STORE_F32_SEQ, METRIC_IP>(rows);
    case 16: return pick_rows<16, STORE_F32_SEQ, METRIC_IP>(rows);
    default: return nullptr;
    }
}

kfn_t gvs_kernel_mw_f32seq_ip(int nch, int rows_per_wave)
{
    switch (nch) {
    case 1: return pick_rows_mw<1, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 2: return pick_rows_mw<2, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 3: return pick_rows_mw<3, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 4: return pick_rows_mw<4, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 6: return pick_rows_mw<6, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 8: return pick_rows_mw<8, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 12: return pick_rows_mw<12, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    case 16: return pick_rows_mw<16, STORE_F32_SEQ, METRIC_IP>(rows_per_wave);
    default: return nullptr;
    }
}

dfn_t gvs_descend_f32seq_ip(int nch)
{
    switch (nch) {
    case 1: return k_descend<1, STORE_F32_SEQ, METRIC_IP>;
    case 2: return k_descend<2, STORE_F32_SEQ, METRIC_IP>;
    case 3: return k_descend<3, STORE_F32_SEQ, METRIC_IP>;
    case 4: return k_descend<4, STORE_F32_SEQ, METRIC_IP>;
    case 6: return k_descend<6, STORE_F32_SEQ, METRIC_IP>;
    case 8: return k_descend<8, STORE_F32_SEQ, METRIC_IP>;
    case 12: return k_descend<12, STORE_F32_SEQ, METRIC_IP>;
    case 16: return k_descend<16, STORE_F32_SEQ, METRIC_IP>;
    default: return nullptr;
    }
}
