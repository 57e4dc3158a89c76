/*
 * CPU oracle for graph vector search: TEST INFRASTRUCTURE ONLY (the checker).
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Every function cites the reference file:line it restates; paths are
 * relative to the reference repository root.  HNSWLIB below means
 * networked_FPGA/CPU_programs_oldest/hnswlib.
 *
 * Tie rule: queue order is the lexicographic (dist, id) order.  hnswlib
 * orders its heaps by dist only (CompareByFirst, HNSWLIB/hnswalg.h:70-75),
 * so with exactly equal distances its order depends on heap internals; the
 * oracle and the HIP kernels both break such ties by node id.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------ */
/* distances                                                           */
/* ------------------------------------------------------------------ */

/* HNSWLIB/space_l2.h:6-20 (L2Sqr) and HNSWLIB/space_ip.h:7-19
 * (InnerProductDistance = 1 - sum a_i b_i), scalar order, no contraction. */
float orc_dist_seq(const float *a, const float *b, int dim, int metric)
{
    float res = 0.0f;
    if (metric == ORC_L2) {
        for (int i = 0; i < dim; i++) {
            float t = a[i] - b[i];
            res += t * t;
        }
        return res;
    }
    for (int i = 0; i < dim; i++) res += a[i] * b[i];
    return 1.0f - res;
}

/* The HIP kernels' order: the row is split into chunks of E elements (one
 * 16-byte load: E = 4 fp32, 8 bf16, 16 fp8); lane position p (0..7) of an
 * 8-lane row group accumulates chunks p, p+8, ... with one fma per element
 * in element order, then the 8 partials are combined by an xor butterfly
 * (xor 4, then 2, then 1):
 *     ((s0+s4)+(s2+s6)) + ((s1+s5)+(s3+s7)).
 * a is the (decoded) stored row, b the query; dim_pad is a multiple of 8E. */
static inline float wave_combine(const float s[8])
{
    float a04 = s[0] + s[4], a26 = s[2] + s[6], a15 = s[1] + s[5], a37 = s[3] + s[7];
    return (a04 + a26) + (a15 + a37);
}

float orc_dist_wave_e(const float *a, const float *b, int dim_pad, int metric, int E)
{
    float s[8];
    int nchunk = dim_pad / E;
    for (int p = 0; p < 8; p++) {
        float acc = 0.0f;
        for (int c = p; c < nchunk; c += 8)
            for (int t = 0; t < E; t++) {
                int i = E * c + t;
                if (metric == ORC_L2) {
                    float d = a[i] - b[i];
                    acc = fmaf(d, d, acc);
                } else {
                    acc = fmaf(a[i], b[i], acc);
                }
            }
        s[p] = acc;
    }
    float r = wave_combine(s);
    return metric == ORC_L2 ? r : 1.0f - r;
}

float orc_dist_wave(const float *a, const float *b, int dim_pad, int metric)
{
    return orc_dist_wave_e(a, b, dim_pad, metric, 4);
}

double orc_dist_f64(const float *a, const float *b, int dim, int metric)
{
    double acc = 0.0;
    if (metric == ORC_L2) {
        for (int i = 0; i < dim; i++) {
            double t = (double)a[i] - (double)b[i];
            acc += t * t;
        }
        return acc;
    }
    for (int i = 0; i < dim; i++) acc += (double)a[i] * (double)b[i];
    return 1.0 - acc;
}

/* --- low-precision storage decode (exact) ------------------------- */
static inline float f32_from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t bits_from_f32(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float orc_bf16_to_f32(uint16_t h) { return f32_from_bits((uint32_t)h << 16); }

/* OCP FP8 E4M3 (bias 7, no infinities, max 448): value = (-1)^s 2^(e-7) (1+m/8),
 * subnormal (e == 0) = (-1)^s m 2^-9. */
float orc_e4m3_to_f32(uint8_t v)
{
    uint32_t s = (uint32_t)(v >> 7) << 31;
    uint32_t e = (v >> 3) & 15u, m = v & 7u;
    if (e == 0) {
        float f = (float)m * 0x1p-9f;
        return s ? -f : f;
    }
    return f32_from_bits(s | ((e + 120u) << 23) | (m << 20));
}

uint16_t orc_f32_to_bf16(float f)
{
    uint32_t u = bits_from_f32(f);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;              /* round to nearest even (finite inputs) */
    return (uint16_t)(u >> 16);
}

/* round to nearest even onto the E4M3 grid, saturating at +-448 */
uint8_t orc_f32_to_e4m3(float f)
{
    uint8_t s = signbit(f) ? 0x80 : 0;
    float a = fabsf(f);
    if (!(a == a)) return 0x7F;      /* NaN */
    if (a >= 448.0f) return s | 0x7E;
    /* candidates: enumerate the 127 non-negative finite codes; pick nearest,
     * ties to the even mantissa (code with lsb 0). */
    int best = 0;
    float bestd = a;                  /* code 0 -> 0.0 */
    for (int c = 1; c <= 0x7E; c++) {
        float v = orc_e4m3_to_f32((uint8_t)c);
        float d = fabsf(v - a);
        if (d < bestd || (d == bestd && (c & 1) == 0 && (best & 1) == 1)) {
            bestd = d;
            best = c;
        }
    }
    return s | (uint8_t)best;
}

static inline float stored_elem(const orc_graph *g, int64_t id, int i)
{
    int64_t off = id * (int64_t)g->dim_pad + i;
    switch (g->storage) {
    case ORC_STORE_BF16: return orc_bf16_to_f32(g->vec_bf16[off]);
    case ORC_STORE_FP8: return orc_e4m3_to_f32(g->vec_fp8[off]) * g->vec_scale[id];
    default: return g->vec_f32[off];
    }
}

/* FP8 rows in the kernels' order: as orc_dist_wave_e with E = 16, where the
 * element term of L2 is d = fma(code, scale, -q) (one rounding), then
 * acc = fma(d, d, acc), and for IP acc = fma(code, q, acc) with each lane's
 * partial multiplied by the scale once before the butterfly. */
static float dist_wave_fp8(const uint8_t *codes, float scale, const float *q, int dim_pad, int metric)
{
    float s[8];
    int nchunk = dim_pad / 16;
    for (int p = 0; p < 8; p++) {
        float acc = 0.0f;
        for (int c = p; c < nchunk; c += 8)
            for (int t = 0; t < 16; t++) {
                int i = 16 * c + t;
                float x = orc_e4m3_to_f32(codes[i]);
                if (metric == ORC_L2) {
                    float d = fmaf(x, scale, -q[i]);
                    acc = fmaf(d, d, acc);
                } else {
                    acc = fmaf(x, q[i], acc);
                }
            }
        s[p] = metric == ORC_L2 ? acc : acc * scale;
    }
    float r = wave_combine(s);
    return metric == ORC_L2 ? r : 1.0f - r;
}

float orc_dist(const orc_graph *g, const float *q, int64_t id, int order)
{
    if (g->storage == ORC_STORE_FP8 && order == ORC_ORDER_WAVE)
        return dist_wave_fp8(g->vec_fp8 + id * (int64_t)g->dim_pad, g->vec_scale[id], q, g->dim_pad, g->metric);
    if (g->storage == ORC_STORE_F32) {
        const float *x = g->vec_f32 + id * (int64_t)g->dim_pad;
        return order == ORC_ORDER_SEQ ? orc_dist_seq(q, x, g->dim, g->metric)
                                      : orc_dist_wave(x, q, g->dim_pad, g->metric);
    }
    /* decode the row once, then use the same summation orders */
    float row[4096];
    int dp = g->dim_pad;
    if (dp > 4096) return NAN;
    for (int i = 0; i < dp; i++) row[i] = stored_elem(g, id, i);
    int E = g->storage == ORC_STORE_BF16 ? 8 : 16;
    return order == ORC_ORDER_SEQ ? orc_dist_seq(q, row, g->dim, g->metric)
                                  : orc_dist_wave_e(row, q, dp, g->metric, E);
}

/* ------------------------------------------------------------------ */
/* (dist, id) heaps                                                     */
/* ------------------------------------------------------------------ */
typedef struct { float d; int32_t id; } item_t;

static inline int key_lt(item_t a, item_t b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

typedef struct { item_t *v; int n, cap; int is_max; } heap_t;

static void heap_init(heap_t *h, int cap, int is_max)
{
    h->v = (item_t *)malloc(sizeof(item_t) * (size_t)(cap > 16 ? cap : 16));
    h->n = 0;
    h->cap = cap > 16 ? cap : 16;
    h->is_max = is_max;
}
static void heap_free(heap_t *h) { free(h->v); }
/* "above" = closer to the top of the heap */
static inline int heap_above(const heap_t *h, item_t a, item_t b)
{
    return h->is_max ? key_lt(b, a) : key_lt(a, b);
}
static void heap_push(heap_t *h, item_t x)
{
    if (h->n == h->cap) {
        h->cap *= 2;
        h->v = (item_t *)realloc(h->v, sizeof(item_t) * (size_t)h->cap);
    }
    int i = h->n++;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!heap_above(h, x, h->v[p])) break;
        h->v[i] = h->v[p];
        i = p;
    }
    h->v[i] = x;
}
static item_t heap_pop(heap_t *h)
{
    item_t top = h->v[0];
    item_t x = h->v[--h->n];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, b = i;
        item_t bv = x;
        if (l < h->n && heap_above(h, h->v[l], bv)) { b = l; bv = h->v[l]; }
        if (r < h->n && heap_above(h, h->v[r], bv)) { b = r; bv = h->v[r]; }
        if (b == i) break;
        h->v[i] = h->v[b];
        i = b;
    }
    if (h->n > 0) h->v[i] = x;
    return top;
}

/* ------------------------------------------------------------------ */
/* visited tags (HNSWLIB/visited_list_pool.h:8-79 restated: epoch tag per node) */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t *tag; int64_t n; uint32_t epoch; } visited_t;
static __thread visited_t tls_vis = {0, 0, 0};

static visited_t *visited_get(int64_t n)
{
    visited_t *v = &tls_vis;
    if (v->n != n) {
        free(v->tag);
        v->tag = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
        v->n = n;
        v->epoch = 0;
    }
    v->epoch++;
    if (v->epoch == 0) {
        memset(v->tag, 0, sizeof(uint32_t) * (size_t)n);
        v->epoch = 1;
    }
    return v;
}

/* ------------------------------------------------------------------ */
/* upper-layer greedy descent: HNSWLIB/hnswalg.h:1119-1147              */
/* ------------------------------------------------------------------ */
int32_t orc_descend(const orc_graph *g, const float *q, int order, float *out_dist, orc_stats *st)
{
    int32_t cur = g->entry;
    float curdist = orc_dist(g, q, cur, order);
    if (st) st->up_evals++;
    for (int level = g->max_level; level > 0; level--) {
        int changed = 1;
        while (changed) {
            changed = 0;
            int64_t row = g->upper_off[cur] + (level - 1);
            const int32_t *nb = g->adj_up + row * (int64_t)g->ru;
            if (st) st->up_hops++;
            for (int i = 0; i < g->ru; i++) {
                int32_t c = nb[i];
                if (c < 0) break;
                float d = orc_dist(g, q, c, order);
                if (st) st->up_evals++;
                if (d < curdist) {      /* strictly closer, HNSWLIB/hnswalg.h:1139 */
                    curdist = d;
                    cur = c;
                    changed = 1;
                }
            }
        }
    }
    if (out_dist) *out_dist = curdist;
    return cur;
}

/* ------------------------------------------------------------------ */
/* layer-0 best-first search                                           */
/* ------------------------------------------------------------------ */
typedef struct { item_t *res; int n, cap; } group_t;

static int drain_sorted(heap_t *top, int32_t *out_ids, float *out_dists)
{
    int n = top->n;
    for (int i = n - 1; i >= 0; i--) {     /* max-heap pops largest first */
        item_t x = heap_pop(top);
        out_ids[i] = x.id;
        out_dists[i] = x.d;
    }
    return n;
}

/* hnswlib searchBaseLayerST<false> (HNSWLIB/hnswalg.h:240-326), generalised
 * to DST: groups of up to mc candidates, up to mg groups in flight.  A group
 * visits and evaluates its candidates' neighbours when it is launched
 * (threshold frozen, vector_search_baselines/unused/hnswalg.h:400-431) and
 * its results are merged into the result heap when it completes; groups
 * complete in launch order and the pipeline is refilled after each
 * completion (FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src/scheduler.hpp:108-154).
 * mc = mg = 1 is exactly searchBaseLayerST. */
static int layer0_impl(const orc_graph *g, const float *q, int32_t ep, int ef,
                       int mc, int mg, int order,
                       int32_t *out_ids, float *out_dists, orc_stats *st,
                       int32_t *trace_ids, int32_t *trace_group, int trace_cap, int *ntrace,
                       const int32_t *init, int init_len)
{
    int64_t ngroups = 0;
    int nt = 0;
    if (ef < 1) ef = 1;
    if (mc < 1) mc = 1;
    if (mg < 1) mg = 1;
    visited_t *vis = visited_get(g->n);
    heap_t top, cand;
    heap_init(&top, ef + 1 + (init ? init_len : 0), 1);
    heap_init(&cand, 4 * ef + (init ? init_len : 0), 0);
    int64_t hops = 0, evals = 0;

    if (!init) {
        item_t e = { orc_dist(g, q, ep, order), ep };
        heap_push(&top, e);
        heap_push(&cand, e);
        vis->tag[ep] = vis->epoch;
        evals = 1;
    } else {
        /* an initial pool instead of one entry: every listed id (in range,
         * first occurrence) is marked visited, evaluated and offered to the
         * result set as a completed group would be (ids outside [0, n) are
         * skipped) */
        for (int i = 0; i < init_len; i++) {
            int32_t x = init[i];
            if (x < 0 || x >= g->n || vis->tag[x] == vis->epoch) continue;
            vis->tag[x] = vis->epoch;
            item_t r = { orc_dist(g, q, x, order), x };
            evals++;
            if (top.n < ef || key_lt(r, top.v[0])) {
                heap_push(&cand, r);
                heap_push(&top, r);
                if (top.n > ef) heap_pop(&top);
            }
        }
    }

    group_t *fifo = (group_t *)calloc((size_t)mg, sizeof(group_t));
    for (int i = 0; i < mg; i++) {
        fifo[i].cap = mc * g->r0 + 1;
        fifo[i].res = (item_t *)malloc(sizeof(item_t) * (size_t)fifo[i].cap);
    }
    int head = 0, inflight = 0;

    for (;;) {
        /* refill the pipeline */
        while (inflight < mg) {
            group_t *gr = &fifo[(head + inflight) % mg];
            gr->n = 0;
            int taken = 0;
            while (taken < mc && cand.n > 0) {
                item_t c = cand.v[0];
                /* break if the closest candidate is beyond the result bound,
                 * HNSWLIB/hnswalg.h:267 */
                if (key_lt(top.v[0], c)) break;
                heap_pop(&cand);
                if (nt < trace_cap && trace_ids) {
                    trace_ids[nt] = c.id;
                    if (trace_group) trace_group[nt] = (int32_t)ngroups;
                }
                nt++;
                taken++;
                hops++;
                const int32_t *nb = g->adj0 + (int64_t)c.id * g->r0;
                for (int j = 0; j < g->r0; j++) {
                    int32_t x = nb[j];
                    if (x < 0) break;
                    if (vis->tag[x] == vis->epoch) continue;
                    vis->tag[x] = vis->epoch;
                    item_t r = { orc_dist(g, q, x, order), x };
                    evals++;
                    gr->res[gr->n++] = r;
                }
            }
            if (taken == 0) break;
            ngroups++;
            inflight++;
        }
        if (inflight == 0) break;
        /* complete the oldest group: HNSWLIB/hnswalg.h:303-319 per result */
        group_t *gr = &fifo[head];
        head = (head + 1) % mg;
        inflight--;
        for (int i = 0; i < gr->n; i++) {
            item_t r = gr->res[i];
            if (top.n < ef || key_lt(r, top.v[0])) {
                heap_push(&cand, r);
                heap_push(&top, r);
                if (top.n > ef) heap_pop(&top);
            }
        }
    }
    for (int i = 0; i < mg; i++) free(fifo[i].res);
    free(fifo);
    int n = drain_sorted(&top, out_ids, out_dists);
    heap_free(&top);
    heap_free(&cand);
    if (st) {
        st->hops += hops;
        st->evals += evals;
        st->groups += ngroups;
    }
    if (ntrace) *ntrace = nt;
    return n;
}

int orc_search_layer0(const orc_graph *g, const float *q, int32_t ep, int ef,
                      int mc, int mg, int order,
                      int32_t *out_ids, float *out_dists, orc_stats *st)
{
    return layer0_impl(g, q, ep, ef, mc, mg, order, out_ids, out_dists, st, NULL, NULL, 0, NULL, NULL, 0);
}

/* HNSWLIB/hnswalg.h:1114-1168 searchKnn, in closer-first order
 * (HNSWLIB/hnswlib.h:171-185). */
int orc_search_knn_trace(const orc_graph *g, const float *q, int k, int ef,
                         int mc, int mg, int order, const int32_t *entry_or_null,
                         int32_t *out_ids, float *out_dists, orc_stats *st,
                         int32_t *trace_ids, int32_t *trace_group, int trace_cap, int *ntrace)
{
    if (ntrace) *ntrace = 0;
    if (g->n <= 0 || k <= 0) return 0;
    int32_t ep = entry_or_null ? *entry_or_null : orc_descend(g, q, order, NULL, st);
    int efe = ef > k ? ef : k;
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)efe);
    float *ds = (float *)malloc(sizeof(float) * (size_t)efe);
    int n = layer0_impl(g, q, ep, efe, mc, mg, order, ids, ds, st, trace_ids, trace_group, trace_cap, ntrace,
                        NULL, 0);
    int m = n < k ? n : k;
    memcpy(out_ids, ids, sizeof(int32_t) * (size_t)m);
    memcpy(out_dists, ds, sizeof(float) * (size_t)m);
    for (int i = m; i < k; i++) { out_ids[i] = -1; out_dists[i] = INFINITY; }
    free(ids);
    free(ds);
    return m;
}

int orc_search_knn(const orc_graph *g, const float *q, int k, int ef,
                   int mc, int mg, int order, const int32_t *entry_or_null,
                   int32_t *out_ids, float *out_dists, orc_stats *st)
{
    return orc_search_knn_trace(g, q, k, ef, mc, mg, order, entry_or_null, out_ids, out_dists, st,
                                NULL, NULL, 0, NULL);
}

/* searchKnn from an initial candidate pool instead of the descent: the
 * layer-0 search (DST mc, mg) seeded with init[0..init_len), trimmed to k. */
int orc_search_pool(const orc_graph *g, const float *q, int k, int ef, int mc, int mg, int order,
                    const int32_t *init, int init_len, int32_t *out_ids, float *out_dists, orc_stats *st)
{
    if (g->n <= 0 || k <= 0) return 0;
    int efe = ef > k ? ef : k;
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)efe);
    float *ds = (float *)malloc(sizeof(float) * (size_t)efe);
    int n = layer0_impl(g, q, -1, efe, mc, mg, order, ids, ds, st, NULL, NULL, 0, NULL, init, init_len);
    int m = n < k ? n : k;
    memcpy(out_ids, ids, sizeof(int32_t) * (size_t)m);
    memcpy(out_dists, ds, sizeof(float) * (size_t)m);
    for (int i = m; i < k; i++) { out_ids[i] = -1; out_dists[i] = INFINITY; }
    free(ids);
    free(ds);
    return m;
}

/* IndexNSG::search_with_base_graph, vector_search_baselines/scripts_nsg/nsg.py:62-125,
 * restated on its own data structure: retset, an array of (dist, id, flag)
 * kept sorted, holding at most L + 1 entries (nsg.py:113-115).  The initial
 * ids (nsg.py:66-80: the navigating node's first neighbours, then random
 * ids not yet flagged) are given by the caller; each is flagged, evaluated
 * and entered unexpanded (nsg.py:82-90).  Then k walks the first L
 * positions: an unexpanded entry is expanded, each unflagged neighbour is
 * flagged and evaluated, and kept if below the L-th entry (nsg.py:102-112);
 * k moves back to the smallest position an insertion took, else on by one
 * (nsg.py:117-121).  Order is by (dist, id): nsg.py orders by distance only
 * (insort_left puts a new entry before equal ones, and an entry equal to the
 * L-th is not kept); the two agree when distances are distinct.  With fewer
 * than L entries every new one is kept (nsg.py always starts with L). */
int orc_nsg_search(const orc_graph *g, const float *q, int K, int L, const int32_t *init, int init_len,
                   int order, int32_t *out_ids, float *out_dists, orc_stats *st)
{
    if (L < 1) L = 1;
    visited_t *vis = visited_get(g->n);
    item_t *ret = (item_t *)malloc(sizeof(item_t) * (size_t)(L + 2 + init_len));
    unsigned char *flag = (unsigned char *)malloc((size_t)(L + 2 + init_len));
    int rn = 0;
    int64_t evals = 0, hops = 0;
    for (int i = 0; i < init_len; i++) {
        int32_t x = init[i];
        if (x < 0 || x >= g->n || vis->tag[x] == vis->epoch) continue;
        vis->tag[x] = vis->epoch;
        item_t r = { orc_dist(g, q, x, order), x };
        evals++;
        /* insertion sort keeps the array ordered (nsg.py sorts once) */
        int pos = rn;
        while (pos > 0 && key_lt(r, ret[pos - 1])) {
            ret[pos] = ret[pos - 1];
            flag[pos] = flag[pos - 1];
            pos--;
        }
        ret[pos] = r;
        flag[pos] = 1;
        rn++;
    }
    int k = 0;
    while (k < L) {
        int nk = L;
        if (k < rn && flag[k]) {
            flag[k] = 0;
            int32_t n = ret[k].id;
            hops++;
            const int32_t *nb = g->adj0 + (int64_t)n * g->r0;
            for (int m = 0; m < g->r0; m++) {
                int32_t x = nb[m];
                if (x < 0) break;
                if (vis->tag[x] == vis->epoch) continue;
                vis->tag[x] = vis->epoch;
                item_t r = { orc_dist(g, q, x, order), x };
                evals++;
                if (rn >= L && !key_lt(r, ret[L - 1])) continue;
                int pos = rn;
                while (pos > 0 && key_lt(r, ret[pos - 1])) {
                    ret[pos] = ret[pos - 1];
                    flag[pos] = flag[pos - 1];
                    pos--;
                }
                ret[pos] = r;
                flag[pos] = 1;
                rn++;
                if (rn > L + 1) rn--;
                if (pos < nk) nk = pos;
            }
        }
        if (nk <= k)
            k = nk;
        else
            k++;
    }
    int m = rn < K ? rn : K;
    if (m > L) m = L;
    for (int i = 0; i < K; i++) {
        out_ids[i] = i < m ? ret[i].id : -1;
        out_dists[i] = i < m ? ret[i].d : INFINITY;
    }
    free(ret);
    free(flag);
    if (st) {
        st->hops += hops;
        st->evals += evals;
    }
    return m;
}

int orc_search_batch_ex(const orc_graph *g, const float *q, int64_t nq, int k, int ef,
                        int mc, int mg, int order, const int32_t *entries_or_null,
                        int32_t *out_ids, float *out_dists, int64_t *out_stats, int stats_width,
                        int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (stats_width < 4) stats_width = 4;
    int64_t i;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
    for (i = 0; i < nq; i++) {
        orc_stats st = {0, 0, 0, 0, 0};
        orc_search_knn(g, q + i * (int64_t)g->dim_pad, k, ef, mc, mg, order,
                       entries_or_null ? entries_or_null + i : NULL,
                       out_ids + i * (int64_t)k, out_dists + i * (int64_t)k, &st);
        if (out_stats) {
            int64_t *row = out_stats + (int64_t)stats_width * i;
            row[0] = st.hops;
            row[1] = st.evals;
            row[2] = st.up_hops;
            row[3] = st.up_evals;
            if (stats_width > 4) row[4] = st.groups;
        }
    }
    return 0;
}

int orc_search_batch(const orc_graph *g, const float *q, int64_t nq, int k, int ef,
                     int mc, int mg, int order, const int32_t *entries_or_null,
                     int32_t *out_ids, float *out_dists, int64_t *out_stats4,
                     int nthreads)
{
    return orc_search_batch_ex(g, q, nq, k, ef, mc, mg, order, entries_or_null, out_ids, out_dists,
                               out_stats4, 4, nthreads);
}

/* ------------------------------------------------------------------ */
/* FP64 brute force (HNSWLIB/bruteforce.h semantics, exact arithmetic)  */
/* ------------------------------------------------------------------ */
typedef struct { double d; int32_t id; } ditem_t;
static inline int dkey_lt(ditem_t a, ditem_t b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

void orc_bruteforce_f64(const float *base, int64_t n, int dim, int metric,
                        const float *q, int64_t nq, int k,
                        int32_t *out_ids, double *out_dists, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    int64_t qi;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (qi = 0; qi < nq; qi++) {
        /* bounded max-heap of size k */
        ditem_t *h = (ditem_t *)malloc(sizeof(ditem_t) * (size_t)(k + 1));
        int hn = 0;
        const float *qq = q + qi * dim;
        for (int64_t j = 0; j < n; j++) {
            ditem_t x = { orc_dist_f64(qq, base + j * dim, dim, metric), (int32_t)j };
            if (hn == k && !dkey_lt(x, h[0])) continue;
            if (hn == k) {
                /* replace top and sift down */
                int i = 0;
                for (;;) {
                    int l = 2 * i + 1, r = l + 1, b = i;
                    ditem_t bv = x;
                    if (l < hn && dkey_lt(bv, h[l])) { b = l; bv = h[l]; }
                    if (r < hn && dkey_lt(bv, h[r])) { b = r; bv = h[r]; }
                    if (b == i) break;
                    h[i] = h[b];
                    i = b;
                }
                h[i] = x;
            } else {
                int i = hn++;
                while (i > 0) {
                    int p = (i - 1) / 2;
                    if (!dkey_lt(h[p], x)) break;
                    h[i] = h[p];
                    i = p;
                }
                h[i] = x;
            }
        }
        /* sort ascending */
        for (int a = 1; a < hn; a++) {
            ditem_t x = h[a];
            int b = a - 1;
            while (b >= 0 && dkey_lt(x, h[b])) { h[b + 1] = h[b]; b--; }
            h[b + 1] = x;
        }
        for (int a = 0; a < k; a++) {
            out_ids[qi * k + a] = a < hn ? h[a].id : -1;
            out_dists[qi * k + a] = a < hn ? h[a].d : INFINITY;
        }
        free(h);
    }
}

int orc_rerank_f32(const float *base_f32, int dim_pad, int metric, const float *q,
                   const int32_t *cand, int ncand, int k, int order,
                   int32_t *out_ids, float *out_dists)
{
    item_t *v = (item_t *)malloc(sizeof(item_t) * (size_t)(ncand > 0 ? ncand : 1));
    int n = 0;
    for (int i = 0; i < ncand; i++) {
        if (cand[i] < 0) continue;
        const float *x = base_f32 + (int64_t)cand[i] * dim_pad;
        float d = order == ORC_ORDER_SEQ ? orc_dist_seq(q, x, dim_pad, metric)
                                         : orc_dist_wave(x, q, dim_pad, metric);
        v[n].d = d;
        v[n].id = cand[i];
        n++;
    }
    for (int a = 1; a < n; a++) {
        item_t x = v[a];
        int b = a - 1;
        while (b >= 0 && key_lt(x, v[b])) { v[b + 1] = v[b]; b--; }
        v[b + 1] = x;
    }
    int m = n < k ? n : k;
    for (int a = 0; a < k; a++) {
        out_ids[a] = a < m ? v[a].id : -1;
        out_dists[a] = a < m ? v[a].d : INFINITY;
    }
    free(v);
    return m;
}

/* ------------------------------------------------------------------ */
/* MurmurHash2 (Austin Appleby's published algorithm), as used by      */
/* unit_tests/bloom_and_hash/sw_murmur2/murmur2.cpp:8-90 and            */
/* FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src/bloom_filter.hpp:43-66 */
/* ------------------------------------------------------------------ */
uint32_t orc_murmur2(const void *key, int len, uint32_t seed)
{
    const uint32_t m = 0x5bd1e995u;
    const unsigned char *p = (const unsigned char *)key;
    uint32_t h = seed ^ (uint32_t)len;
    while (len >= 4) {
        uint32_t k;
        memcpy(&k, p, 4);
        k *= m;
        k ^= k >> 24;
        k *= m;
        h *= m;
        h ^= k;
        p += 4;
        len -= 4;
    }
    switch (len) {
    case 3: h ^= (uint32_t)p[2] << 16; /* fall through */
    case 2: h ^= (uint32_t)p[1] << 8;  /* fall through */
    case 1: h ^= p[0]; h *= m;
    }
    h ^= h >> 13;
    h *= m;
    h ^= h >> 15;
    return h;
}

uint32_t orc_murmur2_keylen4(uint32_t key, uint32_t seed)
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

/* Blocked Bloom filter as the HIP kernels keep it in LDS: one 64-bit word
 * per id chosen by hash(id, seed), six bit positions from hash(id, seed + 1)
 * (five 6-bit fields) and the top six bits of hash(id, seed).  The reference
 * uses three hash functions with seeds hash_seed + pe over a bit array
 * (bloom_filter.hpp:174-188, :274-288). */
void orc_bloom_word_mask(uint32_t id, uint32_t seed, uint32_t nwords, uint32_t *word, uint64_t *mask)
{
    uint32_t h0 = orc_murmur2_keylen4(id, seed);
    uint32_t h1 = orc_murmur2_keylen4(id, seed + 1u);
    *word = h0 & (nwords - 1u);
    *mask = (1ull << (h1 & 63u)) | (1ull << ((h1 >> 6) & 63u)) | (1ull << ((h1 >> 12) & 63u)) |
            (1ull << ((h1 >> 18) & 63u)) | (1ull << ((h1 >> 24) & 63u)) | (1ull << (h0 >> 26));
}

int orc_bloom_test_and_set(uint64_t *words, uint32_t nwords, uint32_t seed, uint32_t id)
{
    uint32_t w;
    uint64_t m;
    orc_bloom_word_mask(id, seed, nwords, &w, &m);
    uint64_t old = words[w];
    words[w] = old | m;
    return (old & m) == m;
}
