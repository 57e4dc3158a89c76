/*
 * gvs.h - C ABI of the MI355X graph vector search library (libgvs.so).
 *
 * The library searches an HNSW (or flat NSG-style) graph whose adjacency
 * lists and vectors are resident in HBM, one wavefront per query by default.
 * All entry points are plain C: pointers, sizes and int status codes; no
 * entry point exits the process.  Errors return a negative gvs_status and
 * set a thread-local message readable with gvs_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   INTER = FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src
 *   HNSWLIB = networked_FPGA/CPU_programs_oldest/hnswlib
 */
#ifndef GVS_H
#define GVS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GVS_API __attribute__((visibility("default")))

typedef enum {
    GVS_OK = 0,
    GVS_ERR_ARG = -1,          /* invalid argument */
    GVS_ERR_HIP = -2,          /* HIP runtime error (message has the HIP error string) */
    GVS_ERR_NOMEM = -3,        /* host or device allocation failed */
    GVS_ERR_UNSUPPORTED = -4,  /* shape/mode not compiled in (e.g. dim > 512) */
    GVS_ERR_OVERFLOW = -5,     /* a per-query visited table overflowed after retries */
    GVS_ERR_IO = -6,           /* file could not be read or has a bad format */
    GVS_ERR_STATE = -7         /* call not valid in the object's current state */
} gvs_status;

typedef enum { GVS_METRIC_L2 = 0, GVS_METRIC_IP = 1 } gvs_metric;
typedef enum { GVS_STORE_F32 = 0, GVS_STORE_BF16 = 1, GVS_STORE_FP8 = 2 } gvs_storage;

/* Flat host description of a graph.  Layer 0: n rows of r0 int32 neighbour
 * ids, unused slots = -1.  Upper layers: node v with level L >= 1 owns rows
 * upper_off[v] .. upper_off[v]+L-1 of adj_up (row l-1 = level l), each ru
 * int32 ids, unused = -1; upper_off[v] = -1 for level-0 nodes.  This is the
 * information of hnswlib's level-0 block and linkLists_
 * (HNSWLIB/hnswalg.h:44-66, :371-381) and of the FPGA layout's
 * ground_links / upper_links files (vector_search_baselines/scripts_hnsw/hnsw.py:264-377). */
typedef struct gvs_graph {
    int64_t n;
    int32_t dim;
    int32_t metric;            /* gvs_metric */
    const float *vectors;      /* n rows */
    int64_t ld_vectors;        /* row stride in floats, >= dim */
    int32_t r0;                /* layer-0 row width (maxM0) */
    const int32_t *adj0;       /* n * r0 */
    int32_t max_level;         /* 0 for a flat graph */
    int32_t entry;             /* enterpoint_node_ (or NSG navigating node) */
    int32_t ru;                /* upper row width (maxM) */
    int64_t up_rows;           /* rows in adj_up */
    const int64_t *upper_off;  /* n, may be NULL when max_level == 0 */
    const int32_t *adj_up;     /* up_rows * ru, may be NULL when max_level == 0 */
    const int64_t *labels;     /* n external labels, NULL = identity */
} gvs_graph;

/* Search parameters.  k, ef as in hnswlib (ef_eff = max(ef, k),
 * HNSWLIB/hnswalg.h:1150-1156).  mc/mg as the reference's DST scheduler
 * (max_cand_batch_size / max_async_stage_num, INTER/scheduler.hpp:100-157);
 * mc = mg = 1 is plain best-first search. */
typedef struct gvs_params {
    int32_t k;
    int32_t ef;
    int32_t mc;                /* candidates expanded per step (1..8) */
    int32_t mg;                /* groups in flight per query (1..8) */
    int32_t waves_per_query;   /* 1 = inter-query; 2,4,8 = intra-query (one workgroup per query) */
    int32_t storage;           /* gvs_storage traversed; the copy must exist */
    int32_t rerank;            /* 1: exact fp32 rerank of the final ef candidates */
    int32_t persistent;        /* 1: persistent kernel draining a device work queue */
    int32_t visited_log2;      /* log2 HBM visited-hash slots per query; 0 = auto */
    int32_t bloom_log2;        /* log2 LDS Bloom words per query; 0 = auto */
    int32_t grid_waves;        /* persistent mode: resident waves; 0 = auto */
    int32_t flags;             /* GVS_FLAG_* */
} gvs_params;

/* gvs_params.flags: GVS_FLAG_ASYNC = enqueue only, no host synchronisation.
 * Queries whose visited table overflowed then return label -1; the caller
 * reads the count with gvs_search_overflow_count() after synchronising and
 * re-issues them (the synchronous form retries them itself). */
#define GVS_FLAG_ASYNC 1
/* Visited-set placement.  By default a query with ef_eff <= 128 keeps an exact
 * visited hash in LDS; a query that fills it to 3/4 freezes it and goes on
 * with an exact per-query hash in HBM for the ids it meets later.  ef_eff in
 * (128, 512] keeps an exact LDS hash too (one wave per query): 2^13 4-byte
 * slots where their 32 KiB cost no waves the launch can use, else 2-byte
 * slots (2^14 or 2^13; graphs of up to 2^(slots' log2 + 8) nodes), frozen
 * the same way; larger ef use an LDS Bloom filter in front of the HBM hash
 * from the start.  GVS_FLAG_HBM_VISITED forces the HBM form,
 * GVS_FLAG_LDS_VISITED the LDS form (visited_log2 in [10, 14] then sets the
 * LDS hash size; its slots are 2 bytes when ef_eff > 128 and the graph fits). */
#define GVS_FLAG_HBM_VISITED 2
#define GVS_FLAG_LDS_VISITED 4
/* With GVS_FLAG_LDS_VISITED: 4-byte LDS hash slots (id + 1) at every ef,
 * instead of 2-byte ones above ef 128. */
#define GVS_FLAG_LDS_WIDE 32
/* Region order (persistent mode, no given entries, graphs with >= 2 upper
 * levels): an upper-layer descent kernel first finds every query's entry
 * and the nodes it settled on at levels 3 and 2; the batch is sorted by
 * those (a coarse address of the query's region) and the layer-0 search
 * runs in that order from eight work queues, one per XCD, so queries of
 * one region run together on one XCD and share its L2.  Results and
 * counters are those of the unordered search.  Ignored by captured
 * searches (gvs_search_graph_create). */
#define GVS_FLAG_REGION_ORDER 8
/* hnswlib order (fp32 storage): every distance is summed as hnswlib's scalar
 * L2Sqr / InnerProductDistance does, res += t*t for i = 0..D-1 with no fused
 * multiply-add (HNSWLIB/space_l2.h:6-20, HNSWLIB/space_ip.h:7-19), one lane
 * per row, so ids, distances and counters equal hnswlib's searchKnn bit for
 * bit.  A parity mode: the default 8-lane order is faster and differs from
 * it only in the last bits of each distance. */
#define GVS_FLAG_HNSWLIB_ORDER 16

/* Per-query debug counters written when stats != NULL (4 int32 per query):
 * layer-0 hops, layer-0 distance evaluations, upper-layer list scans,
 * upper-layer distance evaluations.  The reference's mem_debug carries
 * {hops, visited} per query (INTER/vadd.cpp:88-94, INTER/scheduler.hpp:98-166). */
#define GVS_STATS_PER_QUERY 4
/* int32 per query in every stats buffer this library reads or writes:
 * GVS_STATS_PER_QUERY, or 16 in a build with -DGVS_PROFILE (the counters,
 * then per-phase timers).  Size stats buffers from this.  No reference
 * counterpart (mem_debug has a fixed debug_size, INTER/host.cpp:851-995). */
GVS_API int gvs_stats_per_query(void);

typedef struct gvs_index gvs_index;
typedef struct gvs_host_graph gvs_host_graph;

GVS_API const char *gvs_last_error(void);
GVS_API const char *gvs_version(void);

/* ---- host-side graph construction and formats ---------------------- */

/* HNSW build with hnswlib's addPoint semantics (HNSWLIB/hnswalg.h:982-1112:
 * level draw, greedy descent, searchBaseLayer with efConstruction,
 * getNeighborsByHeuristic2, mutuallyConnectNewElement).  nthreads = 1 gives
 * hnswlib's single-thread graph; exact_order = 1 sums distances in hnswlib's
 * scalar order (bit-identical graph), 0 uses a faster order. */
GVS_API int gvs_hnsw_build(const float *x, int64_t n, int32_t dim, int32_t metric,
                           int32_t M, int32_t ef_construction, uint64_t seed,
                           int32_t nthreads, int32_t exact_order, gvs_host_graph **out);
/* HNSW construction on the GPU, with hnswlib's insertion rules (addPoint,
 * HNSWLIB/hnswalg.h:982-1112; the reference builds on the CPU,
 * vector_search_baselines/scripts_hnsw/construct_and_search_hnsw.py:126-148).
 * Nodes go in index order in batches that grow with the graph: each node of
 * a batch is inserted against the graph as it stood at the batch start
 * (layer 0 on the GPU: construction search, heuristic, reverse links grouped
 * by target, re-pruned when full; levels >= 1 on host threads), so a node
 * only misses links to the nodes of its own batch.  The levels, M, maxM0 =
 * 2M and efc = max(ef_construction, M) are hnswlib's; a batch of one node is
 * addPoint, and with exact_order (hnswlib's scalar summation order) and
 * batches of one the graph is hnswlib's single-thread graph bit for bit.
 * Needs 2 <= M <= 32, dim <= 512, efc <= 1024.  stats (may be NULL):
 * seconds in all, in the host levels, in the layer-0 searches, in
 * selection + links; the number of batches. */
typedef struct gvs_build_config {
    int32_t exact_order;       /* 1: hnswlib's scalar summation order throughout */
    int32_t max_batch;         /* largest batch (0 = 32768) */
    int32_t host_threads;      /* threads for the levels >= 1 (0 = 16) */
    int32_t reserved;
    double batch_ratio;        /* batch size / nodes inserted (0 = 0.01; < 0 = batches of one).  On
                                * 1M x 128, M = 32, efc = 200: 0.01 builds in 2.5 s with recall@10 at
                                * ef 58 within 0.001 of the sequential build, 0.05 in 1.6 s, 0.017 lower */
} gvs_build_config;
GVS_API int gvs_hnsw_build_device(const float *x, int64_t n, int32_t dim, int32_t metric, int32_t M,
                                  int32_t ef_construction, uint64_t seed, const gvs_build_config *cfg,
                                  gvs_host_graph **out, double stats[8]);
/* The batched insertion of gvs_hnsw_build_device run entirely on host
 * threads (same batch schedule and rules; a check of the GPU builder: with
 * exact_order both give the same graph).  No reference counterpart beyond
 * addPoint itself (a batch of one node). */
GVS_API int gvs_hnsw_build_batched(const float *x, int64_t n, int32_t dim, int32_t metric, int32_t M,
                                   int32_t ef_construction, uint64_t seed, int32_t nthreads, int32_t exact_order,
                                   double batch_ratio, int32_t max_batch, gvs_host_graph **out);
/* NSG-style flat graph from an HNSW graph: one row of <= R neighbours per
 * node and a single navigating entry node, the form the FPGA path searches
 * (vector_search_baselines/scripts_nsg/nsg_to_FPGA.py:56-137: fixed-width
 * rows, one entry point).  A node of HNSW level >= 1 keeps up to R/2 of its
 * upper-layer links as long-range edges; the rest of each row is its
 * layer-0 list, pruned with the occlusion rule when it does not fit.  The
 * navigating node is the node nearest to the dataset centroid (among nodes
 * of level >= 1 when `from` has upper layers). */
GVS_API int gvs_nsg_from_hnsw(const gvs_host_graph *from, int32_t R, gvs_host_graph **out);
/* hnswlib .bin reader / writer (HNSWLIB/hnswalg.h:584-700). */
GVS_API int gvs_hnswlib_load(const char *path, int32_t dim, int32_t metric, gvs_host_graph **out);
GVS_API int gvs_hnswlib_save(const gvs_host_graph *g, const char *path);
/* NSG .nsg reader (width, ep, then per node k, ids[k];
 * vector_search_baselines/scripts_nsg/nsg_to_FPGA.py:31-47); vectors are
 * attached from x (n x dim, copied). */
GVS_API int gvs_nsg_load(const char *path, const float *x, int64_t n, int32_t dim,
                         int32_t metric, gvs_host_graph **out);
/* The reference's FPGA-layout index directory (the files the FPGA host
 * loads, INTER/host.cpp:203-224, :537-632), written by
 * vector_search_baselines/scripts_hnsw/hnsw.py:264-377 (kind 0, HNSW: meta.bin
 * = n, maxlevel, entry, maxM, maxM0; upper_links.bin + upper_links_pointers.bin;
 * ground_labels.bin) and scripts_nsg/nsg_to_FPGA.py:56-137 (kind 1, NSG:
 * meta.bin = n, entry, width): per node a 64-byte-aligned link record and a
 * vector record with the FPGA's visited pad, split round-robin over
 * nchannels files ground_{links,vectors}_<nchannels>_chan_<c>.bin.  Save
 * writes kind 0 for graphs with upper layers, kind 1 otherwise. */
GVS_API int gvs_fpga_load(const char *dir, int32_t kind, int32_t dim, int32_t nchannels, int32_t metric,
                          gvs_host_graph **out);
GVS_API int gvs_fpga_save(const gvs_host_graph *g, const char *dir, int32_t nchannels);
GVS_API int gvs_host_graph_view(const gvs_host_graph *g, gvs_graph *view);
GVS_API void gvs_host_graph_free(gvs_host_graph *g);

/* Seeded Gaussian mixture: ncenters centres ~ N(0, center_std^2 I) drawn from
 * seed_centers, point i = centre[u_i] + N(0, noise_std^2 I) from seed_points;
 * normalize = 1 scales rows to unit L2 norm.  Deterministic for any nthreads. */
GVS_API int gvs_gen_mixture(uint64_t seed_centers, uint64_t seed_points, int64_t n, int32_t dim,
                            int32_t ncenters, float center_std, float noise_std,
                            int32_t normalize, float *out, int32_t nthreads);

/* ---- device index --------------------------------------------------- */

/* Copy the graph to the current HIP device and keep it resident.
 * storage_mask: bit GVS_STORE_* set = build that vector copy on the device
 * (F32 is always kept; BF16 / FP8-e4m3 with per-vector scale are optional). */
GVS_API int gvs_index_create(const gvs_graph *g, uint32_t storage_mask, gvs_index **out);
GVS_API int gvs_index_destroy(gvs_index *ix);
/* A second handle on the same device index: it shares the resident arrays
 * (no copy) and owns its own search scratch (work-queue counters, statuses,
 * visited tables, region-order buffers).  Searches through one handle share
 * its scratch and must not overlap; searches through different handles may
 * run at the same time on different streams, so that one batch's descent,
 * sort and ramp run under the previous batch's drain.  Destroy the forks
 * before the index (gvs_index_destroy returns GVS_ERR_STATE while forks are
 * open); a fork of a fork is a fork of the first index.  No counterpart in
 * the reference: its host runs one batch at a time on the FPGA kernel
 * (INTER/host.cpp:851-995). */
GVS_API int gvs_index_fork(gvs_index *ix, gvs_index **out);
/* device pointers of the resident arrays (for tests / introspection) */
GVS_API int gvs_index_device_bytes(const gvs_index *ix, int64_t *graph_bytes, int64_t *vector_bytes);

/* Batch searchKnn over nq queries held in DEVICE memory (row stride ldq
 * floats), asynchronous on `stream` (hipStream_t, NULL = default stream).
 * Replaces the FPGA kernel `vadd` (INTER/vadd.cpp:11-95: query_num, ef, mc,
 * mg, entry_point_ids, query_vectors, out_id/out_dist, mem_debug) and
 * hnswlib searchKnn (HNSWLIB/hnswalg.h:1114-1168).  entries (device, nq) =
 * per-query layer-0 entry point (skips the upper-layer descent, as
 * INTER/host.cpp:711-713 does) or NULL.  Outputs (device): labels[nq*k],
 * dists[nq*k] ascending (closer first, HNSWLIB/hnswlib.h:171-185); missing
 * results are label -1 / dist +inf.  stats: see GVS_STATS_PER_QUERY.
 * Searches through one index handle share its scratch and must not overlap
 * (one stream at a time); see gvs_index_fork for concurrent batches. */
GVS_API int gvs_search_device(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                              const int32_t *entries, const gvs_params *p,
                              int64_t *labels, float *dists, int32_t *stats, void *stream);
/* gvs_search_device with no allocation at call time, for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The search runs as the persistent kernel
 * (persistent is implied, region order and waves_per_query > 1 are not
 * available); its per-query status words and per-wave visited tables live in
 * a fixed 64 MiB device array of the library, one per GPU, held until the
 * process ends; the grid is capped to the visited tables that fit beside the
 * status words, and a batch whose status words leave no room for one
 * workgroup's tables returns GVS_ERR_ARG.  Like GVS_FLAG_ASYNC, queries whose
 * visited table overflows return label -1 and are counted by
 * gvs_search_overflow_count(); there is no retry.  All calls on one GPU
 * share the array: calls that overlap on one GPU (two streams, two threads)
 * race and must be serialised by the caller. */
GVS_API int gvs_search_device_noalloc(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                      const int32_t *entries, const gvs_params *p,
                                      int64_t *labels, float *dists, int32_t *stats, void *stream);
/* Number of queries of the last search on this index whose visited table
 * overflowed (only non-zero after a GVS_FLAG_ASYNC search; synchronises). */
GVS_API int gvs_search_overflow_count(gvs_index *ix, int64_t *count);
/* Device counters of the last search on this index (synchronises):
 * [0] queries claimed from the persistent work queue, [1] queries whose
 * visited table overflowed, [2] queries that outgrew the LDS visited hash and
 * went on with the HBM table.  No reference counterpart (diagnostics). */
GVS_API int gvs_search_counters(gvs_index *ix, int64_t counters[3]);
/* Device-to-device copy through a plain streaming kernel (16-byte aligned
 * pointers and size).  No reference counterpart; used to calibrate HBM byte
 * counters against a kernel of known traffic. */
GVS_API int gvs_device_copy(void *dst, const void *src, int64_t bytes, void *stream);
/* The launch plan gvs_search_device would use for nq queries with *p:
 * plan[0] log2 LDS visited-hash slots (0 = none: Bloom filter + HBM table
 * from the start), [1] 1 if those slots are 2 bytes, [2] log2 HBM
 * visited-table words per wave, [3] log2 LDS Bloom words, [4] LDS bytes per
 * wave (per workgroup with several waves per query), [5] waves per
 * workgroup, [6] resident waves, [7] candidate-list capacity.  No reference
 * counterpart (diagnostics). */
GVS_API int gvs_search_plan(gvs_index *ix, int64_t nq, const gvs_params *p, int32_t plan[8]);
/* MurmurHash2_KeyLen4(keys[i], seeds[i]) -> out[i] for n DEVICE words, with
 * the device function the LDS Bloom filter hashes ids with; the reference's
 * Bloom hash (INTER/bloom_filter.hpp:43-66,
 * unit_tests/bloom_and_hash/sw_murmur2/murmur2.cpp:61-90).  For checking the
 * device hash against known answers. */
GVS_API int gvs_hash_murmur2_keylen4_device(const uint32_t *keys, const uint32_t *seeds, uint32_t *out,
                                            int64_t n, void *stream);
/* Autotuner: times the search of nq DEVICE queries (k, ef, storage, rerank
 * and flags from *p; labels/dists are device result buffers it may overwrite)
 * over mc in {1,2,4,8} (as r0 allows) x mg in {1,2,4} x persistent in {0,1}
 * (persistent also at half the resident waves), and for nq <= 64 also
 * waves_per_query in {2,4,8} x mg in {1,2,4}; above ef_eff 128, unless p->flags fixes the
 * visited placement, each one-wave candidate also with GVS_FLAG_HBM_VISITED
 * (best->flags says which won); writes the fastest configuration to *best and
 * its time per launch to *best_ms.  DST settings trade evaluations for steps
 * and can change recall: checking recall is the caller's part.  Replaces the
 * reference's exhaustive (mc, mg) timing, perf_test_scripts/perf_test_throughput.py:83-133. */
GVS_API int gvs_autotune(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq, const gvs_params *p,
                         int64_t *labels, float *dists, void *stream, gvs_params *best, float *best_ms);
/* Captured search: the kernels of one gvs_search_device call (same
 * arguments; the device buffers are fixed at capture time) recorded into a
 * hipGraph, with visited tables of its own.  Each launch replays the search
 * on whatever the query buffer then holds, with one graph launch instead of
 * the per-call planning and kernel launches.  Replays behave like
 * GVS_FLAG_ASYNC: a query whose visited table overflows returns label -1 and
 * is counted by gvs_search_overflow_count().  Launches of one graph must not
 * overlap (they share its tables).  No reference counterpart: the FPGA host
 * re-enqueues its kernel per batch (INTER/host.cpp:960-995). */
typedef struct gvs_search_graph gvs_search_graph;
GVS_API int gvs_search_graph_create(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                    const int32_t *entries, const gvs_params *p, int64_t *labels,
                                    float *dists, int32_t *stats, gvs_search_graph **out);
GVS_API int gvs_search_graph_launch(gvs_search_graph *g, void *stream);
GVS_API int gvs_search_graph_destroy(gvs_search_graph *g);
/* Same with HOST buffers; synchronous (copies in and out). */
GVS_API int gvs_search_host(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                            const int32_t *entries, const gvs_params *p,
                            int64_t *labels, float *dists, int32_t *stats);

/* NSG search from an initial candidate pool (IndexNSG::search_with_base_graph,
 * vector_search_baselines/scripts_nsg/nsg.py:62-125), L_search = max(ef, k):
 * instead of the upper-layer descent from one entry, query i starts from the
 * init_len ids at init_ids + i * init_ld (device buffer; -1 = none, ids
 * outside [0, n) skipped), each marked visited and evaluated, and the search
 * then expands the best unexpanded of the ef best until none is left (with
 * DST mc / mg as p gives them; nsg.py expands one at a time, mc = mg = 1).
 * init_ids NULL: the first min(ef, r0) neighbours of the graph's entry (the
 * navigating node), which is nsg.py's start without the random ids that the
 * reference's NSG fork removed (vector_search_baselines/scripts_nsg/README_nsg.md:5);
 * a caller that wants nsg.py's random fill passes those ids in the pool.
 * Requires 1 <= init_len <= max(ef, k); one wave per query, no entries, no
 * region order.  Otherwise as gvs_search_device. */
GVS_API int gvs_search_pool_device(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                   const int32_t *init_ids, int32_t init_len, int64_t init_ld,
                                   const gvs_params *p, int64_t *labels, float *dists, int32_t *stats,
                                   void *stream);
/* gvs_search_pool_device with host buffers, synchronous. */
GVS_API int gvs_search_pool_host(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                 const int32_t *init_ids, int32_t init_len, int64_t init_ld, const gvs_params *p,
                                 int64_t *labels, float *dists, int32_t *stats);

/* Exact FP64 brute-force kNN on the device (MFMA f64 GEMM + fused top-k):
 * device queries [nq x dim] (stride ldq), out ids (int32) and FP64 dists,
 * ascending, ties by id.  The reference loads ground truth from files
 * (INTER/host.cpp:148-192); this computes it. */
GVS_API int gvs_bruteforce_f64_device(const gvs_index *ix, const float *queries, int64_t nq,
                                      int64_t ldq, int32_t k, int32_t *ids, double *dists,
                                      void *stream);

/* ---- serving loop ------------------------------------------------------ */
/* Continuous batching: requests with their own k and ef are admitted to the
 * device as query slots and LDS/queue memory free up; a persistent kernel
 * serves them.  Semantics of in-flight windows after the reference's
 * networked client (networked_FPGA/CPU_programs/CPU_client.cpp:68-130,
 * :490-620). */
typedef struct gvs_server gvs_server;
typedef struct gvs_server_config {
    int32_t max_inflight;      /* device query slots */
    int32_t max_k;             /* largest k accepted */
    int32_t max_ef;            /* largest ef accepted */
    int32_t queue_bytes;       /* queue memory the in-flight requests may hold, in bytes (0 = only the
                                * slots limit them): a request holds its query (4 dim), its k result
                                * rows (12 k), its counters and control words; one that does not fit
                                * waits in the caller, as for a busy slot */
    int32_t grid_waves;        /* persistent workers, 0 = auto (resident
                                * waves, at most max_inflight); heavy
                                * mixes can saturate higher below that */
    int32_t mc;                /* DST candidates per step used for every request */
    int32_t storage;           /* vector copy traversed: GVS_STORE_F32, or GVS_STORE_BF16 /
                                * GVS_STORE_FP8 (built for the index; one wave per request) */
    int32_t waves_per_request; /* 1 (0) = one wave per request; 2, 4, 8 = one workgroup of that many
                                * waves per request (the intra-query search, DST(mc, 1); lower
                                * latency for large ef) */
    int32_t rerank;            /* BF16 / FP8 storage: 1 = exact fp32 rerank of each request's final
                                * ef candidates, as in gvs_params */
} gvs_server_config;

GVS_API int gvs_server_open(gvs_index *ix, const gvs_server_config *cfg, gvs_server **out);
/* Submits and polls may come from several threads at once: the claims of
 * ring positions and of completion entries are serialised, the request
 * writes and result copies are not, and each poll returns completions no
 * other poll returns. */
/* Enqueue one request (host query of dim floats).  Admission is by memory:
 * the request's ef and k must fit the LDS the serving waves were planned
 * with (ef <= max_ef, k <= max_k), and a slot and its queue memory must be
 * free.  Returns a request id >= 0, GVS_ERR_ARG when k or ef is outside the
 * admitted range (counted as rejected), or GVS_ERR_STATE when all
 * max_inflight device slots are busy or the queue memory is taken (the
 * caller polls and retries). */
GVS_API int64_t gvs_server_submit(gvs_server *s, const float *query, int32_t k, int32_t ef);
/* Enqueue up to n requests (host queries, row stride ldq floats, with their
 * own ks[i] and efs[i]) in order, stopping at the first that finds no free
 * slot, and publish them to the device together (one release of the ring
 * tail instead of one per request).  Writes req_ids[0..m) and returns m >= 0,
 * the number admitted; the caller re-submits the rest after a poll.  Returns
 * GVS_ERR_ARG when the first request's k or ef is outside the admitted
 * range (later ones stop the batch there).  The reference client sends its
 * queries in batches of batch_size (CPU_client.cpp:450-488). */
GVS_API int64_t gvs_server_submit_batch(gvs_server *s, const float *queries, int64_t n, int64_t ldq,
                                        const int32_t *ks, const int32_t *efs, int64_t *req_ids);
/* Copy out up to max_done finished requests: ids[i], labels[i*max_k..],
 * dists[i*max_k..], latency_us[i] (submit -> completion observed on host).
 * Returns the number copied (>= 0). */
GVS_API int gvs_server_poll(gvs_server *s, int32_t max_done, int64_t *req_ids,
                            int64_t *labels, float *dists, double *latency_us);
/* As gvs_server_poll, plus the request's GVS_STATS_PER_QUERY counters. */
GVS_API int gvs_server_poll_stats(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels,
                                  float *dists, double *latency_us, int32_t *stats);
/* As gvs_server_poll_stats (any output may be NULL), plus status[i]: GVS_OK,
 * or GVS_ERR_OVERFLOW when the request's visited table overflowed (its rows
 * are then label -1 / +inf; re-issue it, e.g. on gvs_search_host, which
 * retries with larger tables). */
GVS_API int gvs_server_poll_status(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels,
                                   float *dists, double *latency_us, int32_t *stats, int32_t *status);
/* Cancel a request that has not completed; returns GVS_OK or GVS_ERR_STATE. */
GVS_API int gvs_server_cancel(gvs_server *s, int64_t req_id);
/* counters: submitted, admitted, completed, cancelled, rejected, device_loops */
GVS_API int gvs_server_stats(gvs_server *s, int64_t counters[6]);
/* The first n of: submitted, admitted, completed, cancelled, rejected,
 * device_loops, overflowed (completed with GVS_ERR_OVERFLOW); returns how
 * many were written. */
GVS_API int gvs_server_counters(gvs_server *s, int64_t *counters, int32_t n);
GVS_API int gvs_server_close(gvs_server *s);
/* Open-loop load generator over an open server: query i (host, stride ldq)
 * arrives at a seeded Poisson time of rate rate_qps with its own k = ks[i]
 * and ef = efs[i]; it is submitted once it has arrived and a slot is free.
 * Outputs per query i: labels/dists rows of max_k, latency_us[i] from
 * arrival to completion; wall_s = time to drain all nq.  The reference's
 * client measures per-batch latency the same way around its windows
 * (networked_FPGA/CPU_programs/CPU_client.cpp:742-790).  One submit thread
 * and the calling thread polling; no other thread may poll the server
 * meanwhile. */
/* Host-side breakdown of the last gvs_server_loadgen on this server: [0] ns
 * per request spent in the submit thread's batched submits (lock
 * included), [1] ns per request in the collecting thread's polls that
 * returned completions, [2] and [3] the fraction of the run's wall time each
 * thread spent there, [4] requests per submitted batch, [5] completions per
 * productive poll.  No reference counterpart (diagnostics). */
GVS_API int gvs_server_loadgen_breakdown(gvs_server *s, double out[6]);
GVS_API int gvs_server_loadgen(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps,
                               uint64_t seed, const int32_t *ks, const int32_t *efs, int64_t *labels,
                               float *dists, double *latency_us, double *wall_s);
/* gvs_server_loadgen with submit_threads threads submitting (requests dealt
 * out in runs of 64 consecutive ones, each thread's runs in arrival order)
 * and poll_threads threads polling (the calling thread is one), 1 to 16
 * each: the host side of a server fed by several connections.  No
 * reference counterpart beyond the client's one send and one receive
 * thread (CPU_client.cpp:490-620). */
GVS_API int gvs_server_loadgen_mt(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps,
                                  uint64_t seed, const int32_t *ks, const int32_t *efs, int64_t *out_labels,
                                  float *out_dists, double *out_latency_us, double *out_wall_s,
                                  int32_t submit_threads, int32_t poll_threads);

/* Network front end with the reference's wire format
 * (networked_FPGA/CPU_programs/README.md:40-72): the node side of the
 * two-connection protocol that networked_FPGA/CPU_programs/CPU_client.cpp
 * drives (:450-620), standing where the FPGA network kernel does
 * (networked_FPGA/kernel/user_krnl/FPGA_inter_query_v1_3/src/hls/
 * FPGA_inter_query_v1_3.cpp:64-100, :187-230).  It listens for the client's
 * query connection on c2f_port (0 = any free port, written to *bound_port as
 * soon as it listens), connects to the client's result listener at
 * f2c_host:f2c_port, and serves until a batch header with batch_size <= 0:
 * per batch a 64-byte header (int32 batch_size first), then batch_size
 * queries of ceil(dim/16) 64-byte packets (fp32, zero padded); per query, in
 * arrival order, a 64-byte header (int32 ef first), ceil(ef/16) packets of
 * int32 ids and as many of fp32 distances, ascending.  Queries are searched
 * on the open server s with k = ef as they arrive; one that overflows its
 * visited table is re-run on the batch path.  Ids are the index's labels
 * (its node ids when it was created without labels, which is what the FPGA
 * returns).  stats (may be NULL): batches, queries, replies, re-runs, bytes
 * in, bytes out.  Blocks until the session ends; GVS_ERR_IO on a socket
 * error or a connection closed mid-session. */
typedef struct gvs_net_config {
    int32_t c2f_port;          /* port the client sends queries to (0 = any) */
    const char *f2c_host;      /* client address that receives results */
    int32_t f2c_port;          /* client port that receives results */
    int32_t dim;               /* query dimension on the wire */
    int32_t ef;                /* results per query (<= the server's max_k and max_ef) */
    int32_t timeout_ms;        /* connection set-up limit (0 = 10 s) */
} gvs_net_config;
GVS_API int gvs_net_serve(gvs_server *s, const gvs_net_config *cfg, int32_t *bound_port, int64_t stats[6]);
/* Wire sizes for dim and ef: [0] batch header, [1] one query, [2] one
 * reply (header + ids + distances), in bytes. */
GVS_API int gvs_net_frame_bytes(int32_t dim, int32_t ef, int64_t out[3]);

#ifdef __cplusplus
}
#endif
#endif
