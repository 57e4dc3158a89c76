/*
 * CPU oracle for graph vector search: TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP search path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to compute the expected answer or the CPU baseline timing.  The
 * product library never links it and has no CPU fallback through it.
 *
 * What it restates (reference file:line, paths relative to the reference
 * repository root):
 *   - hnswlib HierarchicalNSW::searchKnn upper-layer greedy descent
 *       networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:1114-1147
 *   - searchBaseLayerST (two-heap best-first search on layer 0)
 *       networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:240-326
 *   - ef_eff = max(ef, k) and the trim to k
 *       networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h:1150-1167
 *   - L2Sqr / InnerProductDistance (1 - dot)
 *       networked_FPGA/CPU_programs_oldest/hnswlib/space_l2.h:6-20,
 *       networked_FPGA/CPU_programs_oldest/hnswlib/space_ip.h:7-19
 *   - delayed-synchronization traversal (DST): up to mc candidates per
 *     group (vector_search_baselines/unused/hnswalg.h:343-485) and up to mg
 *     groups in flight, refilled when the oldest group completes
 *       FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src/scheduler.hpp:100-157
 *   - MurmurHash2 / MurmurHash2_KeyLen4 and the Bloom filter contract
 *       unit_tests/bloom_and_hash/sw_murmur2/murmur2.cpp:8-90,
 *       unit_tests/bloom_and_hash/sw_bloom/bloom_murmur2.cpp:24-34
 *   - FP64 brute-force kNN (ground truth), as BruteforceSearch
 *       networked_FPGA/CPU_programs_oldest/hnswlib/bruteforce.h
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_dist_seq         == hnswlib scalar L2Sqr / InnerProductDistance
 *                           (bit-exact vs oracle/_ref, and vs FP64 closed form)
 *   orc_dist_wave        within 1e-5 relative of the FP64 closed form
 *   orc_search_knn       vs oracle/_ref (vendored hnswlib compiled from the
 *                           reference sources) and committed golden vectors;
 *                           ef >= N gives the exact kNN (brute force)
 *   orc_search_dst       mc=mg=1 equals orc_search_knn exactly; recall and
 *                           evaluation-count invariants
 *   orc_murmur2*         KeyLen4 == generic MurmurHash2 on 4 bytes; published
 *                           known-answer values
 *   orc_bruteforce_f64   numpy FP64 brute force
 */
#ifndef GVS_ORACLE_H
#define GVS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_L2 = 0, ORC_IP = 1 };
/* distance summation order */
enum { ORC_ORDER_SEQ = 0,   /* hnswlib scalar order: res += t*t, i = 0..D-1 */
       ORC_ORDER_WAVE = 1 };/* the HIP kernel's order (8 lane partials, xor tree) */
/* vector storage */
enum { ORC_STORE_F32 = 0, ORC_STORE_BF16 = 1, ORC_STORE_FP8 = 2 };

typedef struct {
    int64_t n;            /* number of nodes */
    int dim;              /* logical dimension D */
    int dim_pad;          /* stored row length in elements (>= dim; multiple of 32 fp32 / 64 bf16 / 128 fp8 for WAVE order) */
    int metric;           /* ORC_L2 / ORC_IP */
    int storage;          /* ORC_STORE_* */
    const float *vec_f32;     /* [n*dim_pad] when storage == F32 */
    const uint16_t *vec_bf16; /* [n*dim_pad] when storage == BF16 */
    const uint8_t *vec_fp8;   /* [n*dim_pad] e4m3 when storage == FP8 */
    const float *vec_scale;   /* [n] per-vector scale when storage == FP8 */
    int r0;                   /* layer-0 row width; unused slots hold -1 */
    const int32_t *adj0;      /* [n*r0] */
    int max_level;            /* 0 for a flat graph (NSG-style) */
    int32_t entry;            /* global entry point (enterpoint_node_) */
    int ru;                   /* upper-layer row width (M); unused slots hold -1 */
    const int64_t *upper_off; /* [n] index of the node's level-1 row, -1 if level 0 */
    const int32_t *levels;    /* [n] node level */
    const int32_t *adj_up;    /* [rows*ru] rows for levels 1..level, consecutive per node */
} orc_graph;

typedef struct {
    int64_t hops;        /* layer-0 expansions */
    int64_t evals;       /* layer-0 distance evaluations (entry point included) */
    int64_t up_hops;     /* upper-layer neighbour-list scans */
    int64_t up_evals;    /* upper-layer distance evaluations (entry distance included) */
    int64_t groups;      /* layer-0 DST groups launched (== hops when mc == 1) */
} orc_stats;

/* distance between a query (fp32, dim_pad elements, zero padded) and node id */
float orc_dist(const orc_graph *g, const float *q, int64_t id, int order);
/* plain fp32 distance of two rows of length dim in the given order */
float orc_dist_seq(const float *a, const float *b, int dim, int metric);
float orc_dist_wave(const float *a, const float *b, int dim_pad, int metric);
/* E = elements per 16-byte chunk: 4 (fp32), 8 (bf16), 16 (fp8) */
float orc_dist_wave_e(const float *a, const float *b, int dim_pad, int metric, int E);
double orc_dist_f64(const float *a, const float *b, int dim, int metric);

/* decode helpers (exact): bf16 -> f32, e4m3 -> f32 (no scale) */
float orc_bf16_to_f32(uint16_t h);
float orc_e4m3_to_f32(uint8_t v);
/* round-to-nearest-even encoders used to build the low-precision copies */
uint16_t orc_f32_to_bf16(float f);
uint8_t orc_f32_to_e4m3(float f);

/* greedy descent over levels max_level..1 from g->entry; returns layer-0 entry */
int32_t orc_descend(const orc_graph *g, const float *q, int order, float *out_dist, orc_stats *st);

/*
 * Layer-0 search from ep with ef; writes the ef-best (or fewer) results in
 * ascending (dist, id) order.  Returns the number of results.
 * mc = candidates per group, mg = groups in flight; mc = mg = 1 is hnswlib's
 * searchBaseLayerST.
 */
int orc_search_layer0(const orc_graph *g, const float *q, int32_t ep, int ef,
                      int mc, int mg, int order,
                      int32_t *out_ids, float *out_dists, orc_stats *st);

/* searchKnn: descent, layer-0 search with max(ef,k), trim to k (ascending). */
int orc_search_knn(const orc_graph *g, const float *q, int k, int ef,
                   int mc, int mg, int order, const int32_t *entry_or_null,
                   int32_t *out_ids, float *out_dists, orc_stats *st);

/* batch over nq queries (row stride dim_pad), nthreads OpenMP-free pthreads */
int orc_search_batch(const orc_graph *g, const float *q, int64_t nq, int k, int ef,
                     int mc, int mg, int order, const int32_t *entries_or_null,
                     int32_t *out_ids, float *out_dists, int64_t *out_stats4,
                     int nthreads);

/* searchKnn of one query that also records the layer-0 expansion order:
 * trace_ids[i] = the i-th node expanded (popped from the candidate set),
 * trace_group[i] = the DST group (0, 1, ...) that expanded it; at most
 * trace_cap entries are written, *ntrace = the number of expansions. */
int orc_search_knn_trace(const orc_graph *g, const float *q, int k, int ef,
                         int mc, int mg, int order, const int32_t *entry_or_null,
                         int32_t *out_ids, float *out_dists, orc_stats *st,
                         int32_t *trace_ids, int32_t *trace_group, int trace_cap, int *ntrace);

/* searchKnn from an initial candidate pool (no descent, no single entry):
 * init[0..init_len) ids are evaluated, marked visited and seeded as the
 * result set, then the layer-0 search (DST mc, mg) runs with max(ef, k);
 * the NSG search of vector_search_baselines/scripts_nsg/nsg.py:62-125 at
 * mc = mg = 1. */
int orc_search_pool(const orc_graph *g, const float *q, int k, int ef, int mc, int mg, int order,
                    const int32_t *init, int init_len, int32_t *out_ids, float *out_dists, orc_stats *st);

/* nsg.py:62-125 search_with_base_graph restated on its own sorted retset of
 * at most L + 1 entries (an independent check of orc_search_pool); K
 * results, -1 / inf padded. */
int orc_nsg_search(const orc_graph *g, const float *q, int K, int L, const int32_t *init, int init_len,
                   int order, int32_t *out_ids, float *out_dists, orc_stats *st);

/* orc_search_batch with stats_width int64 per query: 4 as above, or 5 with
 * the DST group count appended */
int orc_search_batch_ex(const orc_graph *g, const float *q, int64_t nq, int k, int ef,
                        int mc, int mg, int order, const int32_t *entries_or_null,
                        int32_t *out_ids, float *out_dists, int64_t *out_stats, int stats_width,
                        int nthreads);

/* FP64 exact kNN: out ids/dists (ascending, ties by id) for nq queries */
void orc_bruteforce_f64(const float *base, int64_t n, int dim, int metric,
                        const float *q, int64_t nq, int k,
                        int32_t *out_ids, double *out_dists, int nthreads);

/* exact FP32 rerank of candidate ids (ascending (dist,id)), returns count */
int orc_rerank_f32(const float *base_f32, int dim_pad, int metric, const float *q,
                   const int32_t *cand, int ncand, int k, int order,
                   int32_t *out_ids, float *out_dists);

/* hashing / Bloom filter */
uint32_t orc_murmur2(const void *key, int len, uint32_t seed);
uint32_t orc_murmur2_keylen4(uint32_t key, uint32_t seed);
/* blocked Bloom filter (one 64-bit word per key, 6 bits), word array of nwords (pow2) */
void orc_bloom_word_mask(uint32_t id, uint32_t seed, uint32_t nwords, uint32_t *word, uint64_t *mask);
/* returns 1 if the id was (maybe) present, 0 if definitely absent; inserts it */
int orc_bloom_test_and_set(uint64_t *words, uint32_t nwords, uint32_t seed, uint32_t id);

#ifdef __cplusplus
}
#endif
#endif
