// Search kernel template and the per-(storage, metric) instantiation tables.
#pragma once
#include "search_query.hpp"

namespace gvsk {

typedef void (*kfn_t)(KParams);

template <int NCH, int STORE, int METRIC, int MROWS, bool POOL = false>
#ifndef GVS_MIN_WAVES_PER_EU
#define GVS_MIN_WAVES_PER_EU 1
#endif
// waves per SIMD the descent kernel is compiled to allow (its registers)
#ifndef GVS_DESCEND_WAVES_PER_EU
#define GVS_DESCEND_WAVES_PER_EU 1
#endif
__global__ __launch_bounds__(256, GVS_MIN_WAVES_PER_EU) void k_search(KParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wib = threadIdx.x >> 6;
    unsigned char *lds = smem + (size_t)wib * P.lds_per_wave;
    const int64_t gw = (int64_t)blockIdx.x * P.waves_per_block + wib;
    // One call site for the search, so the kernel holds one copy of it.
    // Work sources: region order - the ordered work is cut into 8 segments,
    // one per XCD (so the queries a XCD runs at once share its L2), and a
    // wave drains its XCD's segment, then helps with the others; persistent -
    // a device work queue, each wave takes the next query; otherwise one
    // query per wave.
    const int x0 = (int)(xcc_id() & 7u);
    int seg = 0;
    bool once = true;
    for (;;) {
        int64_t qi;
        if (P.persistent && P.xcd_queues) {
            int64_t item = -1;
            while (seg < 8) {
                const int x = (x0 + seg) & 7;
                const int64_t s0 = P.nq * x / 8, s1 = P.nq * (x + 1) / 8;
                uint32_t w = 0;
                if (lane_id() == 0) w = atomicAdd(&P.counters[4 + x], 1u);
                w = __builtin_amdgcn_readfirstlane(w);
                if (s0 + (int64_t)w < s1) {
                    item = s0 + (int64_t)w;
                    break;
                }
                seg++;
            }
            if (item < 0) break;
            qi = P.q_perm ? (int64_t)P.q_perm[item] : item;
        } else if (P.persistent) {
            uint32_t q = 0;
            if (lane_id() == 0) q = atomicAdd(&P.counters[0], 1u);
            q = __builtin_amdgcn_readfirstlane(q);
            if ((int64_t)q >= P.nq) break;
            qi = P.q_perm ? (int64_t)P.q_perm[q] : (int64_t)q;
        } else {
            if (!once) break;
            once = false;
            qi = P.q_base + gw;
            if (qi >= P.nq) break;
        }
        search_query<NCH, STORE, METRIC, MROWS, POOL>(P, qi, gw, lds);
    }
}

// Upper-layer descent only, one wave per query: the layer-0 entry, the
// upper-layer counters (into out_stats[.., 2..3]) and a region key from the
// nodes the descent settled on at levels 3 and 2, for ordering the batch.
template <int NCH, int STORE, int METRIC>
__global__ __launch_bounds__(256, GVS_DESCEND_WAVES_PER_EU) void k_descend(KParams P, int32_t *ep_out, uint64_t *key_out, int32_t *perm_out)
{
    constexpr int E = StoreTraits<STORE>::E;
    __shared__ uint64_t keys[4][WAVE];
    __shared__ uint32_t ids[4][WAVE];
    const int wib = threadIdx.x >> 6;
    const int64_t qi = (int64_t)blockIdx.x * 4 + wib;
    if (qi >= P.nq) return;
    WaveLds L;
    L.lkeys = keys[wib];
    L.lids = ids[wib];
    const int lane = lane_id(), p = lane & 7;
    const float *Q = P.queries + qi * P.ldq;
    float qf[NCH * E];
#pragma unroll
    for (int i = 0; i < NCH; i++)
#pragma unroll
        for (int t = 0; t < E; t++) {
            const int idx = (p + 8 * i) * E + t;
            qf[i * E + t] = idx < P.dim ? Q[idx] : 0.0f;
        }
    uint32_t ep, n2, n3;
    float epd;
    int up_hops = 0, up_evals = 0;
    find_entry_path<NCH, STORE, METRIC>(P, qi, L, qf, ep, epd, up_hops, up_evals, n2, n3);
    if (lane == 0) {
        ep_out[qi] = (int32_t)ep;
        key_out[qi] = ((uint64_t)n3 << P.key_bits) | (uint64_t)n2;
        perm_out[qi] = (int32_t)qi;
        if (P.out_stats) {
            P.out_stats[qi * GVS_STATS_ROW + 2] = up_hops;
            P.out_stats[qi * GVS_STATS_ROW + 3] = up_evals;
        }
    }
}

typedef void (*dfn_t)(KParams, int32_t *, uint64_t *, int32_t *);

// rows = mc * ceil(r0 / 64) rounded up to 1, 2 or 8; POOL: the kernels
// with the initial-pool start (the NSG search)
template <int NCH, int STORE, int METRIC, bool POOL = false>
kfn_t pick_rows(int rows)
{
    if (rows <= 1) return k_search<NCH, STORE, METRIC, 1, POOL>;
    if (rows <= 2) return k_search<NCH, STORE, METRIC, 2, POOL>;
    return k_search<NCH, STORE, METRIC, 8, POOL>;
}

// One workgroup of P.waves_per_block waves per query (intra-query).
template <int NCH, int STORE, int METRIC, int WROWS>
__global__ __launch_bounds__(512) void k_search_mw(KParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    if (!P.persistent) {
        const int64_t qi = P.q_base + blockIdx.x;
        if (qi >= P.nq) return;
        search_query_mw<NCH, STORE, METRIC, WROWS>(P, qi, blockIdx.x, smem);
        return;
    }
    // Persistent: the first lane takes the next query from the work queues
    // the one-wave kernel uses (region order: its XCD's segment of the
    // order, then the others'), and the workgroup searches it.  Every wave
    // reads the claim after a barrier and leaves on the same value; the
    // search's own barriers separate this read from the next claim.
    __shared__ int64_t claim;
    const int x0 = (int)(xcc_id() & 7u);
    int seg = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int64_t item = -1;
            if (P.xcd_queues) {
                while (seg < 8) {
                    const int x = (x0 + seg) & 7;
                    const int64_t s0 = P.nq * x / 8, s1 = P.nq * (x + 1) / 8;
                    const uint32_t w = atomicAdd(&P.counters[4 + x], 1u);
                    if (s0 + (int64_t)w < s1) {
                        item = s0 + (int64_t)w;
                        break;
                    }
                    seg++;
                }
            } else {
                const uint32_t w = atomicAdd(&P.counters[0], 1u);
                if ((int64_t)w < P.nq) item = (int64_t)w;
            }
            claim = item < 0 ? -1 : (P.q_perm ? (int64_t)P.q_perm[item] : item);
        }
        wg_sync();
        const int64_t qi = claim;
        if (qi < 0) break;
        search_query_mw<NCH, STORE, METRIC, WROWS>(P, qi, blockIdx.x, smem);
    }
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
// upper-layer descent (region order)
gvsk::kfn_t gvs_kernel_pool_f32_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_pool_f32_ip(int nch, int rows);
gvsk::kfn_t gvs_kernel_pool_bf16_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_pool_bf16_ip(int nch, int rows);
gvsk::kfn_t gvs_kernel_pool_fp8_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_pool_fp8_ip(int nch, int rows);

gvsk::dfn_t gvs_descend_f32_l2(int nch);
gvsk::dfn_t gvs_descend_f32_ip(int nch);
gvsk::dfn_t gvs_descend_bf16_l2(int nch);
gvsk::dfn_t gvs_descend_bf16_ip(int nch);
gvsk::dfn_t gvs_descend_fp8_l2(int nch);
gvsk::dfn_t gvs_descend_fp8_ip(int nch);
// several waves per query (fp32 storage)
gvsk::kfn_t gvs_kernel_mw_f32_l2(int nch, int rows_per_wave);
gvsk::kfn_t gvs_kernel_mw_f32_ip(int nch, int rows_per_wave);
// fp32 rows in hnswlib's scalar summation order (parity mode)
gvsk::kfn_t gvs_kernel_f32seq_l2(int nch, int rows);
gvsk::kfn_t gvs_kernel_f32seq_ip(int nch, int rows);
gvsk::kfn_t gvs_kernel_mw_f32seq_l2(int nch, int rows_per_wave);
gvsk::kfn_t gvs_kernel_mw_f32seq_ip(int nch, int rows_per_wave);
gvsk::dfn_t gvs_descend_f32seq_l2(int nch);
gvsk::dfn_t gvs_descend_f32seq_ip(int nch);
