// Search kernel template and the per-(storage, metric) instantiation tables.
#pragma once
#include "search_query.hpp"

namespace gvsk {

typedef void (*kfn_t)(KParams);

template <int NCH, int STORE, int METRIC, int MROWS>
#ifndef GVS_MIN_WAVES_PER_EU
#define GVS_MIN_WAVES_PER_EU 1
#endif
__global__ __launch_bounds__(256, GVS_MIN_WAVES_PER_EU) void k_search(KParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wib = threadIdx.x >> 6;
    unsigned char *lds = smem + (size_t)wib * P.lds_per_wave;
    const int64_t gw = (int64_t)blockIdx.x * P.waves_per_block + wib;
    if (P.persistent) {
        // device work queue: each wave takes the next query index
        for (;;) {
            uint32_t q = 0;
            if (lane_id() == 0) q = atomicAdd(&P.counters[0], 1u);
            q = __builtin_amdgcn_readfirstlane(q);
            if ((int64_t)q >= P.nq) break;
            search_query<NCH, STORE, METRIC, MROWS>(P, (int64_t)q, gw, lds);
        }
    } else {
        const int64_t qi = P.q_base + gw;
        if (qi >= P.nq) return;
        search_query<NCH, STORE, METRIC, MROWS>(P, qi, gw, lds);
    }
}

// rows = mc * ceil(r0 / 64) rounded up to 1, 2 or 8
template <int NCH, int STORE, int METRIC>
kfn_t pick_rows(int rows)
{
    if (rows <= 1) return k_search<NCH, STORE, METRIC, 1>;
    if (rows <= 2) return k_search<NCH, STORE, METRIC, 2>;
    return k_search<NCH, STORE, METRIC, 8>;
}

// One workgroup of P.waves_per_block waves per query (intra-query).
template <int NCH, int STORE, int METRIC, int WROWS>
__global__ __launch_bounds__(512) void k_search_mw(KParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t qi = P.q_base + blockIdx.x;
    if (qi >= P.nq) return;
    search_query_mw<NCH, STORE, METRIC, WROWS>(P, qi, blockIdx.x, smem);
}

// rows per wave = ceil(mc * ceil(r0/64) / waves) rounded up to 1, 2 or 4
template <int NCH, int STORE, int METRIC>
kfn_t pick_rows_mw(int rows_per_wave)
{
    if (rows_per_wave <= 1) return k_search_mw<NCH, STORE, METRIC, 1>;
    if (rows_per_wave <= 2) return k_search_mw<NCH, STORE, METRIC, 2>;
    if (rows_per_wave <= 4) return k_search_mw<NCH, STORE, METRIC, 4>;
    return nullptr;
}

}  // namespace gvsk

// one table per translation unit; returns nullptr when not compiled
gvsk::kfn_t gvs_kernel_f32_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_f32_ip(int nch, int rows);
gvsk::kfn_t gvs_kernel_bf16_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_bf16_ip(int nch, int rows);
gvsk::kfn_t gvs_kernel_fp8_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_fp8_ip(int nch, int rows);
// several waves per query (fp32 storage)
gvsk::kfn_t gvs_kernel_mw_f32_l2(int nch, int rows_per_wave);
gvsk::kfn_t gvs_kernel_mw_f32_ip(int nch, int rows_per_wave);
