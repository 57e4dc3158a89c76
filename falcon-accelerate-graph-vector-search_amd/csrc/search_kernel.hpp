// Device code for graph search on gfx950 (CDNA4, wave64).
//
// One wavefront owns one query (inter-query parallelism, the reference's
// per-channel pipelines INTER/per_channel_processing_wrapper.hpp:11-185).
// Per wave, in LDS:
//   pool      ef keys, ascending, double-buffered.  key = ord(dist) << 32 |
//             id << 1 | expanded.  The pool is hnswlib's top_candidates; the
//             unexpanded pool entries are its candidate_set restricted to
//             entries that can still be expanded (searchBaseLayerST,
//             HNSWLIB/hnswalg.h:240-326), so one sorted array replaces both
//             heaps and the FPGA's register queues (INTER/priority_queue.hpp).
//   bloom     blocked Bloom filter, one 32-bit word and 3 bits per id,
//             MurmurHash2_KeyLen4 (INTER/bloom_filter.hpp:43-66, :174-188).
//   list      ids to evaluate this step and their (dist, id) keys.
// In HBM, per query slot: an open-addressing hash of visited ids (exact).
// Every neighbour id is inserted with atomicCAS; the CAS winner is the only
// lane that evaluates the id, so the visited set is exact (the FPGA Bloom
// filter alone skips nodes on false positives).  Ids the Bloom filter has
// never seen are gathered speculatively while their CAS is in flight, which
// takes the hash round trip off the critical path.
//
// Distances: 8 lanes per row, 16 bytes per lane per load (one 128-byte line
// per row per load), 8 rows per pass, 8 passes per 64 rows, partials
// reduced by an xor butterfly across the 8 lanes (see orc_dist_wave_e in
// the oracle for the exact summation order).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gvsk {

constexpr int WAVE = 64;
constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr int MAX_ROWS = 8;        // largest mc * ceil(r0/64) adjacency rows per step
constexpr int MAX_MG = 8;          // largest number of DST groups in flight (4-bit sizes in one word)

// LDS words of one wave's (or one workgroup's) ring of selected groups: mg
// lists of up to 8 ids (at least two lists, 64 bytes, for a single wave).
__host__ __device__ constexpr int gvs_sel_words(int mg) { return 8 * (mg < 2 ? 2 : mg); }
// Control words of a workgroup that runs one query on several waves.
constexpr int WG_CTL_WORDS = 16;
constexpr int STORE_F32 = 0, STORE_BF16 = 1, STORE_FP8 = 2;
// fp32 rows summed in hnswlib's scalar order (one lane per row, res += t*t
// for i = 0..D-1, HNSWLIB/space_l2.h:6-20, space_ip.h:7-19) instead of the
// 8-lane order: the search then returns hnswlib's ids and distances bit for
// bit.  A parity mode, not a storage format (it reads the fp32 copy).
constexpr int STORE_F32_SEQ = 3;
constexpr int METRIC_L2 = 0, METRIC_IP = 1;

// int32 per query in a stats row: hops, evaluations, upper hops, upper
// evaluations; the -DGVS_PROFILE build appends 8 phase timers and 4 spare
// words.  Every stats buffer (device, host, server) is sized from this.
#ifdef GVS_PROFILE
#define GVS_STATS_ROW 16
#else
#define GVS_STATS_ROW 4
#endif

struct KParams {
    const void *vec;           // traversal copy (f32 / bf16 / e4m3)
    const float *vscale;       // e4m3 per-vector scale
    const float *vec_f32;      // fp32 copy (rerank)
    int64_t row_elems;         // stored row length of the traversal copy
    int64_t row_elems_f32;     // stored row length of the fp32 copy
    int dim;
    int r0;
    const int32_t *adj0;
    const int32_t *adj_up;
    const int64_t *upper_off;
    int ru;
    int max_level;
    int entry;
    const int64_t *labels;
    const float *queries;
    int64_t ldq;
    int64_t nq;
    int64_t q_base;            // first query of this launch (non-persistent chunks)
    const int32_t *entries;
    int k, ef, mc, mg;
    int rerank;
    uint32_t *vtab;
    int vlog2;
    int blog2;
    int hlog2;                 // > 0: exact visited hash in LDS (2^hlog2 slots); the HBM table only once it is frozen
    int hnarrow;               // LDS hash slots of 2 bytes (quotient + probe step) instead of 4 (id + 1)
    int hbits;                 // id bits the 2-byte slots scramble (hlog2 <= hbits <= hlog2 + 8)
    uint32_t bloom_seed;
    int ef_cap;                // ef rounded up to 64
    int lcap;                  // list capacity (mc * rows * 64)
    int lds_per_wave;          // bytes
    int waves_per_block;
    int persistent;
    int64_t *out_labels;
    float *out_dists;
    int32_t *out_stats;
    int32_t *status;           // per query: 0 ok, 1 visited-table overflow
    uint32_t *counters;        // [0] work queue head, [1] overflow count
    // per-query k / ef (serving); NULL = use k / ef above
    const int32_t *q_k;
    const int32_t *q_ef;
    const int64_t *q_index;    // serving: slot -> query row; NULL = identity
    // region order (persistent mode): work item w is query q_perm[w]; the
    // entries and the upper-layer counters come from the descent kernel
    const int32_t *q_perm;
    int keep_upper_stats;      // leave out_stats[.., 2..3] as the descent wrote them
    int key_bits;              // region key: bits per node id (level-3 node above level-2 node)
    int xcd_queues;            // 8 work queues (counters[4..11]), one per XCD segment of the order
    int out_k_only;            // write only the query's own k results (rows keep stride k above)
    // Initial candidate pool instead of the descent and one entry (the NSG
    // search's start): init_mode 1 = init_len ids per query at
    // init_ids + query row * init_ld (-1 = none), 2 = the first
    // min(ef, r0) neighbours of `entry`; ids outside [0, n) are skipped.
    int init_mode;
    int init_len;
    int64_t init_ld;
    const int32_t *init_ids;
    int64_t n;
};

__device__ __forceinline__ uint32_t ord_f(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f(uint32_t o)
{
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(u);
}
__device__ __forceinline__ uint64_t mk_key(float d, uint32_t id)
{
    return ((uint64_t)ord_f(d) << 32) | ((uint64_t)id << 1);
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return (uint32_t)(k >> 1) & 0x7FFFFFFFu; }
__device__ __forceinline__ float key_dist(uint64_t k) { return unord_f((uint32_t)(k >> 32)); }

// Orders this wave's LDS accesses across lanes.  A wave's LDS instructions
// execute in order, so only the compiler must be kept from moving them: a
// wavefront-scope fence does that without waiting for global memory (a
// workgroup-scope fence would wait for every outstanding load and atomic).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// XCD (accelerator complex die) this wave runs on: HW_REG_XCC_ID (hwreg 20),
// bits 3:0.  Used only to pick a work queue; any value is correct.
__device__ __forceinline__ uint32_t xcc_id()
{
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

__device__ __forceinline__ int mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    uint32_t lo = __shfl_xor((int)(uint32_t)v, m);
    uint32_t hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o = shfl_xor_u64(v, m);
        v = o < v ? o : v;
    }
    return v;
}
// a value the code keeps equal on every lane, made visibly wave-uniform so
// loops bounded by it branch on scalars instead of exec masks
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// v with lane l replaced by the wave-uniform x: v_writelane_b32 with the
// lane select in M0 (a VALU op may read one SGPR besides M0), after one wait
// state for the M0 write the compiler puts in front of it
__device__ __forceinline__ int writelane_i32(int v, int x, int l)
{
    asm("s_nop 0\n\tv_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(x), "{m0}"(l));
    return v;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l)
{
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t murmur2_k4(uint32_t key, uint32_t seed)
{
    const uint32_t m = 0x5bd1e995u;
    uint32_t k = key * m;
    k ^= k >> 24;
    k *= m;
    uint32_t h = (seed ^ 4u) * m;
    h ^= k;
    h ^= h >> 13;
    h *= m;
    h ^= h >> 15;
    return h;
}

template <int STORE>
struct StoreTraits;
template <>
struct StoreTraits<STORE_F32> {
    static constexpr int E = 4;
    static constexpr int BYTES = 4;
};
template <>
struct StoreTraits<STORE_F32_SEQ> {
    static constexpr int E = 4;
    static constexpr int BYTES = 4;
};
template <>
struct StoreTraits<STORE_BF16> {
    static constexpr int E = 8;
    static constexpr int BYTES = 2;
};
template <>
struct StoreTraits<STORE_FP8> {
    static constexpr int E = 16;
    static constexpr int BYTES = 1;
};

// decode one 16-byte chunk into E floats
template <int STORE>
__device__ __forceinline__ void decode_chunk(const uint4 raw, float scale, float *x)
{
    if constexpr (STORE == STORE_F32 || STORE == STORE_F32_SEQ) {
        x[0] = __uint_as_float(raw.x);
        x[1] = __uint_as_float(raw.y);
        x[2] = __uint_as_float(raw.z);
        x[3] = __uint_as_float(raw.w);
    } else if constexpr (STORE == STORE_BF16) {
        const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            x[2 * i] = __uint_as_float(w[i] << 16);
            x[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
        }
    } else {
        // gfx950 converts two OCP e4m3 codes (one 16-bit half of a word) to
        // two fp32 per instruction, exactly.  The per-vector scale is not
        // applied here: elem() folds it into the distance arithmetic.
        const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[i], false);
            const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[i], true);
            x[4 * i + 0] = lo[0];
            x[4 * i + 1] = lo[1];
            x[4 * i + 2] = hi[0];
            x[4 * i + 3] = hi[1];
        }
        (void)scale;
    }
}

// One element's term of a lane's partial.  L2: (x - q)^2, accumulated with
// one fma; for FP8 rows x = code * scale and the difference is formed in one
// rounding, fma(code, scale, -q).  IP: x * q accumulated; FP8 rows accumulate
// code * q and the lane's partial is multiplied by the scale once
// (finish_partial).
template <int STORE, int METRIC>
__device__ __forceinline__ float elem(float x, float q, float scale, float acc)
{
    if constexpr (METRIC == METRIC_L2) {
        const float d = (STORE == STORE_FP8) ? __builtin_fmaf(x, scale, -q) : x - q;
        return __builtin_fmaf(d, d, acc);
    } else {
        (void)scale;
        return __builtin_fmaf(x, q, acc);
    }
}
template <int STORE, int METRIC>
__device__ __forceinline__ float finish_partial(float acc, float scale)
{
    if constexpr (STORE == STORE_FP8 && METRIC == METRIC_IP) return __fmul_rn(acc, scale);
    (void)scale;
    return acc;
}

// Cross-lane moves inside 8-lane groups as DPP operand moves (VALU, no LDS
// round trip): xor 1 and xor 2 are quad permutes; xor 4 is a row shift by 4
// towards the lower half (lanes 0-3 of a group read lane + 4: row_shl) or the
// upper half (lanes 4-7 read lane - 4: row_shr).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float xor1_f(float v) { return dpp_f<0xB1>(v); }  // quad_perm [1,0,3,2]
__device__ __forceinline__ float xor2_f(float v) { return dpp_f<0x4E>(v); }  // quad_perm [2,3,0,1]
__device__ __forceinline__ float xor4_f(float v, bool upper)
{
    const float dn = dpp_f<0x104>(v);  // row_shl:4, lane i reads lane i + 4
    const float up = dpp_f<0x114>(v);  // row_shr:4, lane i reads lane i - 4
    return upper ? up : dn;
}

// 8-lane reduction of 4 passes: in acc[t] = this lane's partial of pass t.
// xor 4 adds the partner's partials (no halving), xor 2 and xor 1 halve, so
// lane p ends with the full sum of pass (p & 3) (lanes p and p^4 agree).
// Per row the order is ((s0+s4)+(s2+s6)) + ((s1+s5)+(s3+s7)) up to operand
// swaps, which IEEE addition does not see.
__device__ __forceinline__ float reduce8x4(const float acc[4], int p)
{
    const bool b2 = p & 4, b1 = p & 2, b0 = p & 1;
    float v4[4], v2[2];
#pragma unroll
    for (int t = 0; t < 4; t++) v4[t] = acc[t] + xor4_f(acc[t], b2);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        float send = b1 ? v4[t] : v4[t + 2];
        float keep = b1 ? v4[t + 2] : v4[t];
        v2[t] = keep + xor2_f(send);
    }
    float send = b0 ? v2[0] : v2[1];
    float keep = b0 ? v2[1] : v2[0];
    return keep + xor1_f(send);
}

// One block of NP passes (8*NP rows from b0).  Every lane loads its chunks
// unconditionally - rows past n repeat row n-1 - so all NP*NCH loads are in
// flight before the first use; rows past n are dropped afterwards.
template <int NCH, int STORE, int METRIC, int NP>
__device__ __forceinline__ void eval_block(const KParams &P, const float *qf, const uint32_t *lids, int b0, int n,
                                           uint64_t *keys)
{
    constexpr int E = StoreTraits<STORE>::E;
    constexpr int B = StoreTraits<STORE>::BYTES;
    const int lane = lane_id();
    const int p = lane & 7, g = lane >> 3;
    uint4 raw[NP][NCH];
    float scale[NP];
#pragma unroll
    for (int t = 0; t < NP; t++) {
        const int r = min(b0 + t * 8 + g, n - 1);
        const uint32_t id = lids[r];
        const unsigned char *row = (const unsigned char *)P.vec + (uint64_t)id * (uint32_t)(P.row_elems * B);
#pragma unroll
        for (int i = 0; i < NCH; i++) raw[t][i] = *(const uint4 *)(row + (size_t)(p + 8 * i) * 16);
        scale[t] = 1.0f;
        if constexpr (STORE == STORE_FP8) scale[t] = P.vscale[id];
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < NP; t++) {
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            float x[E];
            decode_chunk<STORE>(raw[t][i], scale[t], x);
#pragma unroll
            for (int e = 0; e < E; e++) a = elem<STORE, METRIC>(x[e], qf[i * E + e], scale[t], a);
        }
        acc[t] = finish_partial<STORE, METRIC>(a, scale[t]);  // rows past n only meet lanes of the same row, and are dropped below
    }
    const float v = reduce8x4(acc, p);
    const int r = b0 + (p & 3) * 8 + g;
    if (p < 4 && (p & 3) < NP && r < n) {
        const float d = (METRIC == METRIC_L2) ? v : 1.0f - v;
        keys[r] = mk_key(d, lids[r]);
    }
}

// hnswlib's scalar order: lane r sums row lids[b0 + r] element by element,
// res = res + t*t (two roundings, no fma) for L2, res = res + a*b then
// 1 - res for IP.  Element j of the query sits in lane (j/4) % 8 of every
// 8-lane group, register (j/4/8)*4 + j%4, and is broadcast by readlane.
// Padding elements are zero in row and query and add +0.
template <int NCH, int METRIC>
__device__ __forceinline__ void eval_rows_seq(const KParams &P, const float *qf, const uint32_t *lids, int n,
                                              uint64_t *keys)
{
    const int lane = lane_id();
    for (int b0 = 0; b0 < n; b0 += WAVE) {
        const int r = min(b0 + lane, n - 1);
        const uint32_t id = lids[r];
        const float *row = (const float *)P.vec + (uint64_t)id * (uint32_t)P.row_elems;
        float res = 0.0f;
#pragma unroll
        for (int i = 0; i < NCH; i++)
#pragma unroll
            for (int p = 0; p < 8; p++) {
                const float4 x = *(const float4 *)(row + (p + 8 * i) * 4);
                const float xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const float qv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qf[i * 4 + t]), p));
                    if constexpr (METRIC == METRIC_L2) {
                        const float d = __fsub_rn(xv[t], qv);
                        res = __fadd_rn(res, __fmul_rn(d, d));
                    } else {
                        res = __fadd_rn(res, __fmul_rn(xv[t], qv));
                    }
                }
            }
        if (b0 + lane < n) {
            const float d = (METRIC == METRIC_L2) ? res : __fsub_rn(1.0f, res);
            keys[b0 + lane] = mk_key(d, id);
        }
    }
}

// Distances of rows lids[0..n) -> keys[0..n).  All 64 lanes take part;
// rows go in blocks of up to 32 (passes of 8 rows, 8 lanes per row).
template <int NCH, int STORE, int METRIC>
__device__ __forceinline__ void eval_rows(const KParams &P, const float *qf, const uint32_t *lids, int n,
                                          uint64_t *keys)
{
    n = uni(n);
    if constexpr (STORE == STORE_F32_SEQ) {
        eval_rows_seq<NCH, METRIC>(P, qf, lids, n, keys);
        return;
    }
    for (int b0 = 0; b0 < n; b0 += 32) {
        const int nb = n - b0;
        if (nb > 16) eval_block<NCH, STORE, METRIC, 4>(P, qf, lids, b0, n, keys);
        else if (nb > 8) eval_block<NCH, STORE, METRIC, 2>(P, qf, lids, b0, n, keys);
        else eval_block<NCH, STORE, METRIC, 1>(P, qf, lids, b0, n, keys);
    }
}

}  // namespace gvsk
