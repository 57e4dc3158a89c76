// Serving kernel instantiations: fp32 storage, L2 metric.
#include "serve_kernel.hpp"

using namespace gvsk;

sfn_t gvs_serve_kernel_f32_l2(int nch, int rows)
{
    switch (nch) {
    case 1: return pick_serve_rows<1, STORE_F32, METRIC_L2>(rows);
    case 2: return pick_serve_rows<2, STORE_F32, METRIC_L2>(rows);
    case 3: return pick_serve_rows<3, STORE_F32, METRIC_L2>(rows);
    case 4: return pick_serve_rows<4, STORE_F32, METRIC_L2>(rows);
    case 6: return pick_serve_rows<6, STORE_F32, METRIC_L2>(rows);
    case 8: return pick_serve_rows<8, STORE_F32, METRIC_L2>(rows);
    case 12: return pick_serve_rows<12, STORE_F32, METRIC_L2>(rows);
    case 16: return pick_serve_rows<16, STORE_F32, METRIC_L2>(rows);
    default: return nullptr;
    }
}

sfn_t gvs_serve_kernel_mw_f32_l2(int nch, int rows_per_wave)
{
    switch (nch) {
    case 1: return pick_serve_rows_mw<1, STORE_F32, METRIC_L2>(rows_per_wave);
    case 2: return pick_serve_rows_mw<2, STORE_F32, METRIC_L2>(rows_per_wave);
    case 3: return pick_serve_rows_mw<3, STORE_F32, METRIC_L2>(rows_per_wave);
    case 4: return pick_serve_rows_mw<4, STORE_F32, METRIC_L2>(rows_per_wave);
    case 6: return pick_serve_rows_mw<6, STORE_F32, METRIC_L2>(rows_per_wave);
    case 8: return pick_serve_rows_mw<8, STORE_F32, METRIC_L2>(rows_per_wave);
    case 12: return pick_serve_rows_mw<12, STORE_F32, METRIC_L2>(rows_per_wave);
    case 16: return pick_serve_rows_mw<16, STORE_F32, METRIC_L2>(rows_per_wave);
    default: return nullptr;
    }
}
