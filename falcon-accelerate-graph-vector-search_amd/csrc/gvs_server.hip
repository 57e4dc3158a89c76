// Serving loop: continuous batching over a persistent kernel.
//
// Semantics follow the reference's networked client windows
// (networked_FPGA/CPU_programs/CPU_client.cpp:68-130: at most
// query_window_size requests in flight; :490-620 submit / receive threads;
// :742-790 per-request latency), with requests admitted one by one as device
// slots free up instead of in fixed batches, each with its own k and ef.
//
// Host and device share pinned (fine-grained) host memory: the host writes a
// request into a free slot (query, k, ef, sequence number), appends the slot
// to the ring and bumps `tail`; device waves claim ring entries with a CAS on
// a device counter, search, write the results into the slot's pinned result
// rows and store the sequence number into done[slot].  The host polls the
// done flags of its in-flight slots.  Device waves leave the loop on the stop
// flag or after an idle timeout; the host relaunches the kernel when work is
// pending and the previous launch has finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "gvs_internal.hpp"
#include "serve_kernel.hpp"

using namespace gvsk;

namespace gvs {
int plan_search(const gvs_index *ix, const gvs_params *p, int max_ef, int max_k, LaunchPlan *lp,
                int64_t max_waves, bool pool = false);
int ensure_scratch(gvs_index *ix, int64_t slots, int vlog2, int64_t nstatus);
void fill_params(const gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, KParams &P);
}  // namespace gvs

namespace {
struct Slot {
    int64_t req = -1;
    uint64_t ring_idx = 0;
    double t_arrival = 0.0;  // microseconds, steady clock
    int32_t k = 0;           // the request's k (its results are k entries)
    int64_t tag = -1;        // the load generator's query index
    bool busy = false;
    bool cancelled = false;
};

double now_us()
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

struct gvs_server {
    gvs_index *ix = nullptr;
    gvs_server_config cfg{};
    gvs::LaunchPlan lp;
    ServeArgs args{};
    sfn_t fn = nullptr;
    hipStream_t stream = nullptr;
    int64_t grid_waves = 0;
    size_t shm = 0;          // dynamic LDS per workgroup
    bool launched = false;
    // pinned host memory
    ServeCtl *ctl = nullptr;
    uint32_t *ring = nullptr;
    float *qbuf = nullptr;
    int32_t *qk = nullptr, *qef = nullptr;
    uint64_t *qseq = nullptr, *done = nullptr;
    int64_t *labels = nullptr;
    float *dists = nullptr;
    int32_t *stats = nullptr;
    // device
    uint64_t *d_head = nullptr, *d_served = nullptr, *d_mirror = nullptr;
    uint32_t *d_vtab = nullptr;  // the server's own per-wave HBM visited tables
    int32_t *d_status = nullptr;
    // host bookkeeping.  Three locks: submit_mu serialises the claims of
    // ring positions (ring tails), poll_mu the claims of completion entries
    // (completion ring head), and mu guards what both touch (free slots,
    // slot records, the in-flight list, ring-entry reuse, counters, the
    // device loop's state).  Locks are held for bookkeeping only: request
    // writes and result copies run outside them, so several threads can
    // submit and poll at once.
    std::mutex mu, submit_mu, poll_mu;
    std::vector<Slot> slots;
    std::vector<int> free_slots, inflight;
    std::vector<int> inflight_pos;   // slot -> index in inflight
    uint64_t *cring = nullptr;       // pinned completion rings, one per request ring (written by the device)
    uint64_t chead[SERVE_QUEUES] = {};  // poll_mu: completions consumed per ring
    int cnext = 0;                      // poll_mu: ring a poll looks at first
    double last_query_us = -1e30;    // last hipStreamQuery
    std::vector<uint8_t> ring_busy;  // ring entry (q * rcap + i) still names an incomplete request
    struct Admitted { int slot; size_t rpos; int64_t m; uint64_t seq; };
    struct Done { int slot; bool ovf, cancelled; int64_t req, tag; double t_arrival; int k; };
    uint64_t tails[SERVE_QUEUES] = {};                  // submit_mu: ring positions claimed
    std::atomic<uint64_t> published[SERVE_QUEUES]{};    // ... and published to the device, in claim order
    int rr = 0;                              // ring the next request tries first
    uint64_t next_seq = 1;
    int64_t next_req = 0;
    int64_t c_submitted = 0, c_completed = 0, c_cancelled = 0, c_rejected = 0, c_loops = 0, c_overflowed = 0;
    int64_t inflight_bytes = 0;  // mu: queue memory held by admitted requests (config queue_bytes)
    // host-side breakdown of the last gvs_server_loadgen (see gvs.h)
    double lg_breakdown[6] = {0, 0, 0, 0, 0, 0};
};

// Queue memory a request of k results holds while in flight: its query,
// its k result rows (id and distance), its counters and control words.
static int64_t request_bytes(const gvs_server *s, int k)
{
    return 4 * (int64_t)s->ix->dim + 12 * (int64_t)k + 4 * GVS_STATS_ROW + 16;
}

static void serve_free(gvs_server *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamDestroy(s->stream);
    (void)hipHostFree(s->ctl);
    (void)hipHostFree(s->ring);
    (void)hipHostFree(s->qbuf);
    (void)hipHostFree(s->qk);
    (void)hipHostFree(s->qef);
    (void)hipHostFree(s->qseq);
    (void)hipHostFree(s->done);
    (void)hipHostFree(s->cring);
    (void)hipHostFree(s->labels);
    (void)hipHostFree(s->dists);
    (void)hipHostFree(s->stats);
    (void)hipFree(s->d_head);
    (void)hipFree(s->d_mirror);
    (void)hipFree(s->d_served);
    (void)hipFree(s->d_vtab);
    (void)hipFree(s->d_status);
    delete s;
}

// (re)launch the device loop if it is not running; caller holds s->mu
static int serve_ensure_running(gvs_server *s)
{
    if (s->launched) {
        // the loop leaves only after 200 ms without work: looking at the
        // stream every 100 us is enough and keeps the runtime off the path
        const double t = now_us();
        if (t - s->last_query_us < 100.0) return GVS_OK;
        s->last_query_us = t;
        hipError_t q = hipStreamQuery(s->stream);
        if (q == hipErrorNotReady) return GVS_OK;
        if (q != hipSuccess) {
            gvs::set_error("serving kernel failed: %s", hipGetErrorString(q));
            return GVS_ERR_HIP;
        }
        s->launched = false;
    }
    hipLaunchKernelGGL(s->fn, dim3((unsigned)(s->grid_waves / s->lp.wpb)), dim3(s->lp.wpb * WAVE), s->shm, s->stream,
                       s->args);
    if (!gvs::hip_ok(hipGetLastError(), "launch k_serve")) return GVS_ERR_HIP;
    s->launched = true;
    s->c_loops++;
    return GVS_OK;
}

extern "C" int gvs_server_open(gvs_index *ix, const gvs_server_config *cfg, gvs_server **out)
{
    const int W = cfg ? (cfg->waves_per_request > 0 ? cfg->waves_per_request : 1) : 1;
    if (!ix || !cfg || !out || cfg->max_inflight < 1 || cfg->max_inflight > (1 << 20) || cfg->max_k < 1 ||
        cfg->max_ef < 1 || cfg->mc < 0 || cfg->storage < GVS_STORE_F32 || cfg->storage > GVS_STORE_FP8 ||
        (W != 1 && W != 2 && W != 4 && W != 8) || (W > 1 && cfg->storage != GVS_STORE_F32) || cfg->queue_bytes < 0) {
        gvs::set_error("gvs_server_open: invalid configuration (fp32, BF16 or FP8 storage; 1, 2, 4 or 8 waves per "
                       "request, several only with fp32)");
        return GVS_ERR_ARG;
    }
    gvs_server *s = new gvs_server();
    s->ix = ix;
    s->cfg = *cfg;
    if (cfg->queue_bytes > 0 && cfg->queue_bytes < request_bytes(s, cfg->max_k)) {
        gvs::set_error("gvs_server_open: queue_bytes %d cannot hold one request of k = max_k (%lld bytes)",
                       cfg->queue_bytes, (long long)request_bytes(s, cfg->max_k));
        delete s;
        return GVS_ERR_ARG;
    }
    gvs_params p;
    std::memset(&p, 0, sizeof p);
    p.k = cfg->max_k;
    p.ef = cfg->max_ef;
    p.mc = cfg->mc > 0 ? cfg->mc : 1;
    p.mg = 1;
    p.waves_per_query = W;
    p.storage = cfg->storage;
    p.rerank = cfg->storage != GVS_STORE_F32 && cfg->rerank ? 1 : 0;
    p.persistent = 1;
    // the grid holds at most grid_waves (or max_inflight) waves: the
    // visited hash may use LDS that only waves beyond those would need
    int rc = gvs::plan_search(ix, &p, cfg->max_ef, cfg->max_k, &s->lp,
                              cfg->grid_waves > 0 ? cfg->grid_waves : (int64_t)cfg->max_inflight * W);
    if (rc) {
        delete s;
        return rc;
    }
    const int E = cfg->storage == GVS_STORE_F32 ? 4 : cfg->storage == GVS_STORE_BF16 ? 8 : 16;
    const int nch = ix->dim_pad / (8 * E);
    const int rows = p.mc * ((ix->r0 + WAVE - 1) / WAVE);
    const bool l2 = ix->metric == GVS_METRIC_L2;
    if (W == 1) {
        if (cfg->storage == GVS_STORE_F32)
            s->fn = l2 ? gvs_serve_kernel_f32_l2(nch, rows) : gvs_serve_kernel_f32_ip(nch, rows);
        else if (cfg->storage == GVS_STORE_BF16)
            s->fn = l2 ? gvs_serve_kernel_bf16_l2(nch, rows) : gvs_serve_kernel_bf16_ip(nch, rows);
        else
            s->fn = l2 ? gvs_serve_kernel_fp8_l2(nch, rows) : gvs_serve_kernel_fp8_ip(nch, rows);
        s->shm = (size_t)s->lp.wpb * s->lp.lds_per_wave;
    } else {
        // one workgroup of W waves per request (its LDS: lp.lds_per_wave)
        const int rpw = (rows + W - 1) / W;
        s->fn = l2 ? gvs_serve_kernel_mw_f32_l2(nch, rpw) : gvs_serve_kernel_mw_f32_ip(nch, rpw);
        s->shm = (size_t)s->lp.lds_per_wave;
        if (s->fn && s->shm > 65536 &&
            !gvs::hip_ok(hipFuncSetAttribute((const void *)s->fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)s->shm),
                         "hipFuncSetAttribute")) {
            delete s;
            return GVS_ERR_HIP;
        }
    }
    if (!s->fn) {
        delete s;
        gvs::set_error("gvs_server_open: dim %d not compiled for serving", ix->dim);
        return GVS_ERR_UNSUPPORTED;
    }
    int blocks_per_cu = 0;
    if (!gvs::hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, (const void *)s->fn,
                                                                  s->lp.wpb * WAVE, s->shm),
                     "occupancy")) {
        delete s;
        return GVS_ERR_HIP;
    }
    s->grid_waves = cfg->grid_waves > 0 ? cfg->grid_waves
                                        : (int64_t)std::max(1, blocks_per_cu) * ix->cu_count * s->lp.wpb;
    // a request occupies one wave (one workgroup of W), so waves beyond
    // max_inflight requests could never hold work; they would only add pollers
    if (cfg->grid_waves <= 0) s->grid_waves = std::min<int64_t>(s->grid_waves, (int64_t)cfg->max_inflight * W);
    s->grid_waves = (s->grid_waves + s->lp.wpb - 1) / s->lp.wpb * s->lp.wpb;
    const int64_t R = cfg->max_inflight;
    // Each ring holds up to rcap entries; an entry is reused only after the
    // request that held it is complete (checked at admission, which then
    // tries the next ring), so a holder never reads an overwritten entry.
    // Requests go to the rings round robin.
    uint32_t rcap = 16;
    while (rcap < 4u * (uint32_t)((R + SERVE_QUEUES - 1) / SERVE_QUEUES)) rcap <<= 1;
    // completion ring q: at most rcap entries are unconsumed (a request's
    // ring entry is reused only after its completion was polled)
    const uint32_t ccap = 2u * rcap;
    const int dim = ix->dim;
    bool ok = gvs::hip_ok(hipHostMalloc((void **)&s->ctl, sizeof(ServeCtl), hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->ring, 4 * (size_t)rcap * SERVE_QUEUES, hipHostMallocCoherent),
                          "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qbuf, 4 * (size_t)R * dim, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qk, 4 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qef, 4 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qseq, 8 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->done, 8 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->cring, 8 * (size_t)ccap * SERVE_QUEUES, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->labels, 8 * (size_t)R * cfg->max_k, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->dists, 4 * (size_t)R * cfg->max_k, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->stats, 4 * GVS_STATS_ROW * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipMalloc(&s->d_head, 8 * SERVE_QUEUES * SERVE_HEAD_STRIDE), "hipMalloc") &&
              gvs::hip_ok(hipMalloc(&s->d_served, 8 * SERVE_QUEUES * SERVE_HEAD_STRIDE), "hipMalloc") &&
              gvs::hip_ok(hipMalloc(&s->d_mirror, 8 * SERVE_MIRROR_WORDS), "hipMalloc") &&
              gvs::hip_ok(hipMemset(s->d_mirror, 0, 8 * SERVE_MIRROR_WORDS), "hipMemset") &&
              gvs::hip_ok(hipMemset(s->d_head, 0, 8 * SERVE_QUEUES * SERVE_HEAD_STRIDE), "hipMemset") &&
              gvs::hip_ok(hipMemset(s->d_served, 0, 8 * SERVE_QUEUES * SERVE_HEAD_STRIDE), "hipMemset") &&
              gvs::hip_ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    const size_t vwords = (size_t)s->grid_waves << s->lp.vlog2;
    if (ok)
        ok = gvs::hip_ok(hipMalloc(&s->d_vtab, 4 * vwords), "hipMalloc visited tables") &&
             gvs::hip_ok(hipMemset(s->d_vtab, 0, 4 * vwords), "hipMemset visited tables") &&
             gvs::hip_ok(hipMalloc(&s->d_status, 4 * (size_t)R), "hipMalloc status");
    if (!ok) {
        serve_free(s);
        return GVS_ERR_HIP;
    }
    std::memset(s->ctl, 0, sizeof(ServeCtl));
    std::memset(s->done, 0, 8 * (size_t)R);
    std::memset(s->cring, 0, 8 * (size_t)ccap * SERVE_QUEUES);
    s->inflight_pos.assign((size_t)R, -1);
    std::memset(s->qseq, 0, 8 * (size_t)R);
    s->slots.assign((size_t)R, Slot());
    s->ring_busy.assign((size_t)rcap * SERVE_QUEUES, 0);
    for (int64_t i = R - 1; i >= 0; i--) s->free_slots.push_back((int)i);
    KParams P;
    gvs::fill_params(ix, &p, s->lp, P);
    P.k = cfg->max_k;
    P.ef = cfg->max_ef;
    P.ldq = dim;
    P.out_labels = s->labels;
    P.out_dists = s->dists;
    P.out_stats = s->stats;
    P.out_k_only = 1;  // only a request's own k results cross the bus
    P.persistent = 1;
    P.vtab = s->d_vtab;
    P.status = s->d_status;
    s->args.P = P;
    s->args.ctl = s->ctl;
    s->args.ring = s->ring;
    s->args.rcap = rcap;
    s->args.dim = dim;
    s->args.qbuf = s->qbuf;
    s->args.qk = s->qk;
    s->args.qef = s->qef;
    s->args.qseq = s->qseq;
    s->args.done = s->done;
    s->args.cring = s->cring;
    s->args.ccap = ccap;
    s->args.nslots = (uint32_t)R;
    s->args.head = s->d_head;
    s->args.mirror = s->d_mirror;
    s->args.served = s->d_served;
    s->args.idle_ticks = 20 * 1000 * 1000;  // 200 ms at 100 MHz
    std::lock_guard<std::mutex> lk(s->mu);
    int rc2 = serve_ensure_running(s);
    if (rc2) {
        serve_free(s);
        return rc2;
    }
    *out = s;
    return GVS_OK;
}

// Admit up to n requests in order, stopping at the first that is outside
// the admitted range (GVS_ERR_ARG, counted as rejected) or finds no free
// slot or ring entry (GVS_ERR_STATE), write them into their slots and ring
// entries, and publish them together (one release of each ring's tail).
// Claims of ring positions are serialised (submit_mu); the writes are not,
// so several threads may submit at once.  A ring's tail is published in
// claim order: a batch waits until the batches that claimed earlier
// positions of that ring have published theirs.  No batch waits on a later
// one (claims are taken under one lock, so of two batches one is first on
// every ring), and a waiting batch holds no lock.  tags (may be NULL) are
// kept with the requests for the load generator.  Returns the number
// admitted, or the error of the first request when none was.
static int64_t submit_batch_impl(gvs_server *s, const float *queries, int64_t n, int64_t ldq, const int32_t *ks,
                                 const int32_t *efs, int64_t *req_ids, const double *t_arrival,
                                 const int64_t *tags = nullptr)
{
    const uint64_t rmask = (uint64_t)s->args.rcap - 1u;
    int64_t err = 0;
    thread_local std::vector<gvs_server::Admitted> adm;
    adm.clear();
    uint64_t first[SERVE_QUEUES], last[SERVE_QUEUES];  // ring positions this batch claimed: [first, last)
    {
        std::lock_guard<std::mutex> lk(s->submit_mu);
        for (int q = 0; q < SERVE_QUEUES; q++) first[q] = s->tails[q];
        const double t = now_us();
        std::lock_guard<std::mutex> lk2(s->mu);
        for (int64_t m = 0; m < n; m++) {
            const int32_t k = ks[m], ef = efs[m];
            if (k < 1 || k > s->cfg.max_k || ef < 1 || ef > s->cfg.max_ef) {
                s->c_rejected++;
                gvs::set_error("gvs_server_submit: k=%d ef=%d outside the admitted range (max_k %d, max_ef %d)", k,
                               ef, s->cfg.max_k, s->cfg.max_ef);
                err = GVS_ERR_ARG;
                break;
            }
            // a ring whose next entry is free (an entry is reused only after
            // the request that held it is complete), round robin
            int q = -1;
            for (int j = 0; j < SERVE_QUEUES && !s->free_slots.empty(); j++) {
                const int c = (s->rr + j) % SERVE_QUEUES;
                if (!s->ring_busy[(size_t)c * s->args.rcap + (s->tails[c] & rmask)]) {
                    q = c;
                    break;
                }
            }
            if (q < 0) {
                gvs::set_error("gvs_server_submit: all %d device slots busy", s->cfg.max_inflight);
                err = GVS_ERR_STATE;
                break;
            }
            const int64_t rb = request_bytes(s, k);
            if (s->cfg.queue_bytes > 0 && s->inflight_bytes + rb > s->cfg.queue_bytes) {
                gvs::set_error("gvs_server_submit: queue memory taken (%lld of %d bytes in flight)",
                               (long long)s->inflight_bytes, s->cfg.queue_bytes);
                err = GVS_ERR_STATE;
                break;
            }
            s->inflight_bytes += rb;
            s->rr = (q + 1) % SERVE_QUEUES;
            const int slot = s->free_slots.back();
            s->free_slots.pop_back();
            const size_t rpos = (size_t)q * s->args.rcap + (size_t)(s->tails[q] & rmask);
            s->tails[q]++;
            Slot &sl = s->slots[(size_t)slot];
            sl.req = s->next_req++;
            sl.t_arrival = t_arrival ? t_arrival[m] : t;
            sl.k = k;
            sl.tag = tags ? tags[m] : -1;
            sl.busy = true;
            sl.cancelled = false;
            sl.ring_idx = rpos;
            s->ring_busy[rpos] = 1;
            s->inflight_pos[(size_t)slot] = (int)s->inflight.size();
            s->inflight.push_back(slot);
            s->c_submitted++;
            if (req_ids) req_ids[m] = sl.req;
            adm.push_back({slot, rpos, m, s->next_seq++});
        }
        for (int q = 0; q < SERVE_QUEUES; q++) last[q] = s->tails[q];
    }
    // the admitted slots and ring entries belong to this batch until published
    const int dim = s->ix->dim;
    for (const auto &a : adm) {
        std::memcpy(s->qbuf + (size_t)a.slot * dim, queries + a.m * ldq, sizeof(float) * dim);
        s->qk[a.slot] = ks[a.m];
        s->qef[a.slot] = efs[a.m];
        s->qseq[a.slot] = a.seq;
        s->ring[a.rpos] = (uint32_t)a.slot;
    }
    if (!adm.empty()) {
        // request fields and ring entries are globally visible before the tails
        std::atomic_thread_fence(std::memory_order_seq_cst);
        for (int q = 0; q < SERVE_QUEUES; q++) {
            if (last[q] == first[q]) continue;
            for (int spin = 0; s->published[q].load(std::memory_order_acquire) != first[q]; spin++)
                if (spin > 64) std::this_thread::yield();
            __atomic_store_n(&s->ctl->tail[q], last[q], __ATOMIC_RELEASE);
            s->published[q].store(last[q], std::memory_order_release);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        const int rc = serve_ensure_running(s);
        if (rc) return rc;
    }
    if (adm.empty() && err == GVS_ERR_ARG) return err;
    return (int64_t)adm.size();  // 0 when no slot was free (the message says so)
}

extern "C" int64_t gvs_server_submit(gvs_server *s, const float *query, int32_t k, int32_t ef)
{
    if (!s || !query) {
        gvs::set_error("gvs_server_submit: invalid argument");
        return GVS_ERR_ARG;
    }
    int64_t id = -1;
    const int64_t r = submit_batch_impl(s, query, 1, s->ix->dim, &k, &ef, &id, nullptr);
    return r == 1 ? id : r == 0 ? GVS_ERR_STATE : r;
}

extern "C" int64_t gvs_server_submit_batch(gvs_server *s, const float *queries, int64_t n, int64_t ldq,
                                           const int32_t *ks, const int32_t *efs, int64_t *req_ids)
{
    if (!s || n < 0 || (n > 0 && (!queries || !ks || !efs)) || ldq < s->ix->dim) {
        gvs::set_error("gvs_server_submit_batch: invalid argument");
        return GVS_ERR_ARG;
    }
    return submit_batch_impl(s, queries, n, ldq, ks, efs, req_ids, nullptr);
}

// completions ahead whose result rows a poll requests early
#ifndef GVS_POLL_PREFETCH
#define GVS_POLL_PREFETCH 6
#endif

// Deliver up to max_done completions, each ring's in the order the device
// finished them, the rings taken round robin.  Claims of completion entries are serialised (poll_mu); the result
// copies are not, so several threads may poll at once, each receiving its
// own completions.  The device wrote each request's own k results into its
// slot's rows; rows are delivered max_k wide, the rest label -1 / dist +inf.
// tags_out (may be NULL) receives the tags the requests were submitted with.
static int poll_impl(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels, float *dists,
                     double *latency_us, int32_t *stats_out, int32_t *status_out, int64_t *tags_out = nullptr)
{
    const int K = s->cfg.max_k;
    const uint64_t cmask = (uint64_t)s->args.ccap - 1u;
    const uint32_t err = __atomic_load_n(&s->ctl->err, __ATOMIC_ACQUIRE);
    if (err) {
        gvs::set_error("gvs_server_poll: the device read an out-of-range slot index from ring entry %u", err - 1u);
        return GVS_ERR_STATE;
    }
    // entry i of a completion ring carries (i + 1) in its high word once it
    // is written, then the overflow bit and the slot
    thread_local std::vector<gvs_server::Done> comp;
    comp.clear();
    {
        std::lock_guard<std::mutex> lk(s->poll_mu);
        for (int j = 0; j < SERVE_QUEUES && (int32_t)comp.size() < max_done; j++) {
            const int q = (s->cnext + j) % SERVE_QUEUES;
            const uint64_t *cr = s->cring + (size_t)q * s->args.ccap;
            while ((int32_t)comp.size() < max_done) {
                const uint64_t h = s->chead[q];
                const uint64_t v = __atomic_load_n(&cr[h & cmask], __ATOMIC_ACQUIRE);
                if ((uint32_t)(v >> 32) != (uint32_t)(h + 1u)) break;
                const int slot = (int)((uint32_t)v & 0x7FFFFFFFu);
                if (slot >= s->cfg.max_inflight) {
                    gvs::set_error("gvs_server_poll: completion ring %d names slot %d of %d", q, slot,
                                   s->cfg.max_inflight);
                    return GVS_ERR_STATE;
                }
                s->chead[q] = h + 1u;
                comp.push_back({slot, ((uint32_t)v >> 31) != 0u, false, 0, -1, 0.0, 0});
            }
        }
        s->cnext = (s->cnext + 1) % SERVE_QUEUES;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto &c : comp) {
            const Slot &sl = s->slots[(size_t)c.slot];
            if (!sl.busy) {
                gvs::set_error("gvs_server_poll: completion ring names slot %d, which is not in flight", c.slot);
                return GVS_ERR_STATE;
            }
            c.req = sl.req;
            c.tag = sl.tag;
            c.t_arrival = sl.t_arrival;
            c.cancelled = sl.cancelled;
            c.k = sl.k;
        }
        if (comp.empty()) {
            if (!s->inflight.empty()) {
                const int rc = serve_ensure_running(s);  // a launch that went idle leaves work behind
                if (rc) return rc;
            }
            return 0;
        }
    }
    // copies, outside the locks: a slot is freed only below
    const double t = now_us();
    int n = 0;
    const size_t nc = comp.size();
    for (size_t j = 0; j < nc; j++) {
        // the device wrote these rows across the bus: the host reads them from
        // memory, so the rows of a few completions ahead are requested early
        if (j + GVS_POLL_PREFETCH < nc) {
            const int ps = comp[j + GVS_POLL_PREFETCH].slot;
            __builtin_prefetch(s->labels + (size_t)ps * K);
            __builtin_prefetch(s->dists + (size_t)ps * K);
        }
        const auto &c = comp[j];
        if (c.cancelled) continue;
        const int k = std::min(K, std::max(1, c.k));
        if (req_ids) req_ids[n] = c.req;
        if (tags_out) tags_out[n] = c.tag;
        if (labels) {
            std::memcpy(labels + (size_t)n * K, s->labels + (size_t)c.slot * K, 8 * (size_t)k);
            std::fill(labels + (size_t)n * K + k, labels + (size_t)(n + 1) * K, (int64_t)-1);
        }
        if (dists) {
            std::memcpy(dists + (size_t)n * K, s->dists + (size_t)c.slot * K, 4 * (size_t)k);
            std::fill(dists + (size_t)n * K + k, dists + (size_t)(n + 1) * K, INFINITY);
        }
        if (latency_us) latency_us[n] = t - c.t_arrival;
        if (stats_out)
            std::memcpy(stats_out + (size_t)n * GVS_STATS_ROW, s->stats + (size_t)c.slot * GVS_STATS_ROW,
                        4 * GVS_STATS_ROW);
        if (status_out) status_out[n] = c.ovf ? GVS_ERR_OVERFLOW : GVS_OK;
        n++;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (const auto &c : comp) {
        Slot &sl = s->slots[(size_t)c.slot];
        if (!c.cancelled) s->c_completed++;
        if (c.ovf) s->c_overflowed++;
        sl.busy = false;
        s->inflight_bytes -= request_bytes(s, sl.k);
        s->ring_busy[sl.ring_idx] = 0;
        s->free_slots.push_back(c.slot);
        const int ip = s->inflight_pos[(size_t)c.slot];
        const int last = s->inflight.back();
        s->inflight[(size_t)ip] = last;
        s->inflight_pos[(size_t)last] = ip;
        s->inflight.pop_back();
        s->inflight_pos[(size_t)c.slot] = -1;
    }
    if (!s->inflight.empty()) {
        const int rc = serve_ensure_running(s);
        if (rc) return rc;
    }
    return n;
}

extern "C" int gvs_server_poll(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels, float *dists,
                               double *latency_us)
{
    if (!s || max_done < 0) {
        gvs::set_error("gvs_server_poll: invalid argument");
        return GVS_ERR_ARG;
    }
    return poll_impl(s, max_done, req_ids, labels, dists, latency_us, nullptr, nullptr);
}

extern "C" int gvs_server_poll_stats(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels,
                                     float *dists, double *latency_us, int32_t *stats)
{
    if (!s || max_done < 0) {
        gvs::set_error("gvs_server_poll_stats: invalid argument");
        return GVS_ERR_ARG;
    }
    return poll_impl(s, max_done, req_ids, labels, dists, latency_us, stats, nullptr);
}

extern "C" int gvs_server_poll_status(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels,
                                      float *dists, double *latency_us, int32_t *stats, int32_t *status)
{
    if (!s || max_done < 0) {
        gvs::set_error("gvs_server_poll_status: invalid argument");
        return GVS_ERR_ARG;
    }
    return poll_impl(s, max_done, req_ids, labels, dists, latency_us, stats, status);
}

extern "C" int gvs_server_cancel(gvs_server *s, int64_t req_id)
{
    if (!s) return GVS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    for (int slot : s->inflight) {
        Slot &sl = s->slots[(size_t)slot];
        if (sl.req == req_id && !sl.cancelled) {
            // the device may already be working on it: its result is dropped
            sl.cancelled = true;
            s->c_cancelled++;
            return GVS_OK;
        }
    }
    gvs::set_error("gvs_server_cancel: request %lld is not in flight", (long long)req_id);
    return GVS_ERR_STATE;
}

extern "C" int gvs_server_stats(gvs_server *s, int64_t counters[6])
{
    if (!s || !counters) return GVS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    counters[0] = s->c_submitted;
    counters[1] = s->c_submitted;  // every accepted request is admitted to a device slot
    counters[2] = s->c_completed;
    counters[3] = s->c_cancelled;
    counters[4] = s->c_rejected;
    counters[5] = s->c_loops;
    return GVS_OK;
}

extern "C" int gvs_server_counters(gvs_server *s, int64_t *counters, int32_t n)
{
    if (!s || !counters || n < 0) return GVS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t c[7] = {s->c_submitted, s->c_submitted, s->c_completed, s->c_cancelled,
                          s->c_rejected,  s->c_loops,     s->c_overflowed};
    for (int i = 0; i < n && i < 7; i++) counters[i] = c[i];
    return std::min<int32_t>(n, 7);
}

namespace gvs {
// The index a server searches, the parameters its requests run with (k and
// ef are per request) and its k / ef limits, so that a request can be re-run
// on the batch path, which retries visited-table overflows with larger
// tables.
int server_search_setup(gvs_server *s, gvs_index **ix, gvs_params *p, int *max_k, int *max_ef)
{
    if (!s || !ix || !p || !max_k || !max_ef) return GVS_ERR_ARG;
    *ix = s->ix;
    *max_k = s->cfg.max_k;
    *max_ef = s->cfg.max_ef;
    std::memset(p, 0, sizeof *p);
    p->mc = s->cfg.mc > 0 ? s->cfg.mc : 1;
    p->mg = 1;
    p->waves_per_query = s->cfg.waves_per_request > 0 ? s->cfg.waves_per_request : 1;
    p->storage = s->cfg.storage;
    p->rerank = s->cfg.storage != GVS_STORE_F32 && s->cfg.rerank ? 1 : 0;
    return GVS_OK;
}
}  // namespace gvs

extern "C" int gvs_server_close(gvs_server *s)
{
    if (!s) return GVS_ERR_ARG;
    {
        // no new claims; batches already claimed finish their publication first
        std::lock_guard<std::mutex> lk(s->submit_mu);
        for (int q = 0; q < SERVE_QUEUES; q++) {
            while (s->published[q].load(std::memory_order_acquire) != s->tails[q]) std::this_thread::yield();
            __atomic_store_n(&s->ctl->tail[q], s->tails[q] | SERVE_STOP, __ATOMIC_RELEASE);
        }
    }
    // the loop leaves within microseconds of seeing stop; a loop that does
    // not is reported and its buffers are left allocated (it may still use them)
    const double t0 = now_us();
    hipError_t e;
    while ((e = hipStreamQuery(s->stream)) == hipErrorNotReady) {
        if (now_us() - t0 > 30e6) {
            gvs::set_error("gvs_server_close: serving loop did not stop within 30 s");
            return GVS_ERR_STATE;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    serve_free(s);
    return gvs::hip_ok(e, "gvs_server_close") ? GVS_OK : GVS_ERR_HIP;
}

// Open-loop load generator: request i arrives at a seeded Poisson time
// (rate_qps).  As the reference client does with its send and receive
// threads (networked_FPGA/CPU_programs/CPU_client.cpp:490-620), submit
// threads submit every request that has arrived, in one batch per pass (one
// publication to the device), as slots allow, and poll threads (the calling
// thread is one) collect completions; a request's latency runs from its
// arrival to the poll that sees it completed.  With several submit threads
// the requests are dealt out in runs of LG_RUN consecutive ones, run r to
// thread r mod nsub, each thread submitting its runs in arrival order.
static constexpr int64_t LG_RUN = 64;

static int loadgen_impl(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps, uint64_t seed,
                        const int32_t *ks, const int32_t *efs, int64_t *out_labels, float *out_dists,
                        double *out_latency_us, double *out_wall_s, int nsub, int npoll)
{
    const int K = s->cfg.max_k;
    std::vector<double> arrival((size_t)nq);
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_qps / 1e6);  // per microsecond
    double t = 0.0;
    for (int64_t i = 0; i < nq; i++) {
        t += gap(rng);
        arrival[(size_t)i] = t;
    }
    std::vector<int64_t> tag((size_t)nq);
    for (int64_t i = 0; i < nq; i++) tag[(size_t)i] = i;
    const double t0 = now_us();
    std::vector<double> t_arr((size_t)nq);
    for (int64_t i = 0; i < nq; i++) t_arr[(size_t)i] = t0 + arrival[(size_t)i];
    std::atomic<int> err{0};
    std::atomic<bool> stop{false};
    std::atomic<int64_t> submitted{0}, done_n{0};
    std::mutex msg_mu;
    std::string msg;
    auto fail = [&](int rc, const char *m) {
        int z = 0;
        if (err.compare_exchange_strong(z, rc)) {
            std::lock_guard<std::mutex> lk(msg_mu);
            msg = m;
        }
        stop.store(true);
    };
    std::vector<double> sub_us((size_t)nsub, 0.0), poll_us((size_t)npoll, 0.0);
    std::vector<int64_t> sub_batches((size_t)nsub, 0), poll_calls((size_t)npoll, 0);
    // submit thread j: runs j, j + nsub, ... of LG_RUN requests, each run
    // in arrival order, everything that has arrived in one batch per pass
    auto submitter = [&](int j) {
        std::vector<int64_t> ids(LG_RUN);
        for (int64_t r0 = (int64_t)j * LG_RUN; r0 < nq && !stop.load(std::memory_order_relaxed);
             r0 += (int64_t)nsub * LG_RUN) {
            const int64_t r1 = std::min(nq, r0 + LG_RUN);
            int64_t next = r0;
            while (next < r1 && !stop.load(std::memory_order_relaxed)) {
                const double now = now_us() - t0;
                int64_t m = 0;
                while (next + m < r1 && arrival[(size_t)(next + m)] <= now) m++;
                if (m == 0) {
                    std::this_thread::yield();
                    continue;
                }
                const double a0 = now_us();
                const int64_t got = submit_batch_impl(s, queries + next * ldq, m, ldq, ks + next, efs + next,
                                                      ids.data(), t_arr.data() + next, tag.data() + next);
                sub_us[(size_t)j] += now_us() - a0;
                if (got < 0) {
                    fail((int)got, gvs_last_error());
                    return;
                }
                if (got > 0) sub_batches[(size_t)j]++;
                next += got;
                submitted.fetch_add(got, std::memory_order_relaxed);
                if (got < m) std::this_thread::yield();  // no free slot: wait for completions
            }
        }
    };
    auto poller = [&](int j) {
        const int batch = 1024;
        std::vector<int64_t> tg(batch), lab((size_t)batch * K);
        std::vector<float> dd((size_t)batch * K);
        std::vector<double> lat(batch);
        double last_progress = now_us();
        while (done_n.load(std::memory_order_relaxed) < nq && !stop.load(std::memory_order_relaxed)) {
            const double a0 = now_us();
            const int n = poll_impl(s, batch, nullptr, lab.data(), dd.data(), lat.data(), nullptr, nullptr, tg.data());
            if (n < 0) {
                fail(n, gvs_last_error());
                return;
            }
            if (n > 0) {
                poll_us[(size_t)j] += now_us() - a0;
                poll_calls[(size_t)j]++;
            }
            for (int i2 = 0; i2 < n; i2++) {
                const int64_t i = tg[(size_t)i2];
                if (i < 0 || i >= nq) continue;
                if (out_labels) std::memcpy(out_labels + i * K, lab.data() + (size_t)i2 * K, 8 * (size_t)K);
                if (out_dists) std::memcpy(out_dists + i * K, dd.data() + (size_t)i2 * K, 4 * (size_t)K);
                if (out_latency_us) out_latency_us[i] = lat[(size_t)i2];
            }
            const double now = now_us();
            if (n > 0) {
                done_n.fetch_add(n, std::memory_order_relaxed);
                last_progress = now;
            } else {
                if (now - last_progress > 30e6) {
                    char m[200];
                    std::snprintf(m, sizeof m,
                                  "gvs_server_loadgen: no request completed for 30 s (%lld of %lld done, %lld submitted)",
                                  (long long)done_n.load(), (long long)nq, (long long)submitted.load());
                    fail(GVS_ERR_STATE, m);
                    return;
                }
                std::this_thread::yield();
            }
        }
    };
    std::vector<std::thread> th;
    for (int j = 0; j < nsub; j++) th.emplace_back(submitter, j);
    for (int j = 1; j < npoll; j++) th.emplace_back(poller, j);
    poller(0);
    stop.store(true);
    for (auto &x : th) x.join();
    if (err.load()) {
        gvs::set_error("%s", msg.c_str());
        return err.load();
    }
    const double wall_us = now_us() - t0;
    if (out_wall_s) *out_wall_s = wall_us / 1e6;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const double n = (double)std::max<int64_t>(nq, 1);
        double su = 0, pu = 0;
        int64_t sb = 0, pc = 0;
        for (int j = 0; j < nsub; j++) su += sub_us[(size_t)j], sb += sub_batches[(size_t)j];
        for (int j = 0; j < npoll; j++) pu += poll_us[(size_t)j], pc += poll_calls[(size_t)j];
        s->lg_breakdown[0] = 1e3 * su / n;
        s->lg_breakdown[1] = 1e3 * pu / n;
        s->lg_breakdown[2] = su / (wall_us * nsub);
        s->lg_breakdown[3] = pu / (wall_us * npoll);
        s->lg_breakdown[4] = (double)nq / (double)std::max<int64_t>(sb, 1);
        s->lg_breakdown[5] = (double)nq / (double)std::max<int64_t>(pc, 1);
    }
    return GVS_OK;
}

extern "C" int gvs_server_loadgen(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps,
                                  uint64_t seed, const int32_t *ks, const int32_t *efs, int64_t *out_labels,
                                  float *out_dists, double *out_latency_us, double *out_wall_s)
{
    return gvs_server_loadgen_mt(s, queries, nq, ldq, rate_qps, seed, ks, efs, out_labels, out_dists, out_latency_us,
                                 out_wall_s, 1, 1);
}

extern "C" int gvs_server_loadgen_mt(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps,
                                     uint64_t seed, const int32_t *ks, const int32_t *efs, int64_t *out_labels,
                                     float *out_dists, double *out_latency_us, double *out_wall_s,
                                     int32_t submit_threads, int32_t poll_threads)
{
    if (!s || !queries || nq < 0 || ldq < s->ix->dim || rate_qps <= 0.0 || !ks || !efs || submit_threads < 1 ||
        submit_threads > 16 || poll_threads < 1 || poll_threads > 16) {
        gvs::set_error("gvs_server_loadgen: invalid argument (1 to 16 submit and poll threads)");
        return GVS_ERR_ARG;
    }
    return loadgen_impl(s, queries, nq, ldq, rate_qps, seed, ks, efs, out_labels, out_dists, out_latency_us,
                        out_wall_s, submit_threads, poll_threads);
}

extern "C" int gvs_server_loadgen_breakdown(gvs_server *s, double out[6])
{
    if (!s || !out) return GVS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    for (int i = 0; i < 6; i++) out[i] = s->lg_breakdown[i];
    return GVS_OK;
}
