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
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "gvs_internal.hpp"
#include "serve_kernel.hpp"

using namespace gvsk;

namespace gvs {
int plan_search(const gvs_index *ix, const gvs_params *p, int max_ef, int max_k, LaunchPlan *lp,
                int64_t max_waves);
int ensure_scratch(gvs_index *ix, int64_t slots, int vlog2, int64_t nstatus);
void fill_params(const gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, KParams &P);
}  // namespace gvs

namespace {
struct Slot {
    int64_t req = -1;
    uint64_t ring_idx = 0;
    double t_arrival = 0.0;  // microseconds, steady clock
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
    uint64_t *d_head = nullptr, *d_served = nullptr;
    uint32_t *d_vtab = nullptr;  // the server's own per-wave HBM visited tables
    int32_t *d_status = nullptr;
    // host bookkeeping
    std::mutex mu;
    std::vector<Slot> slots;
    std::vector<int> free_slots, inflight;
    std::vector<int> inflight_pos;   // slot -> index in inflight
    uint64_t *cring = nullptr;       // pinned completion ring (written by the device)
    uint64_t chead = 0;              // completions consumed
    double last_query_us = -1e30;    // last hipStreamQuery
    std::vector<uint8_t> ring_busy;  // ring entry still names an incomplete request
    uint64_t next_seq = 1, tail = 0;
    int64_t next_req = 0;
    int64_t c_submitted = 0, c_completed = 0, c_cancelled = 0, c_rejected = 0, c_loops = 0;
};

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
    const size_t shm = (size_t)s->lp.wpb * s->lp.lds_per_wave;
    hipLaunchKernelGGL(s->fn, dim3((unsigned)(s->grid_waves / s->lp.wpb)), dim3(s->lp.wpb * WAVE), shm, s->stream,
                       s->args);
    if (!gvs::hip_ok(hipGetLastError(), "launch k_serve")) return GVS_ERR_HIP;
    s->launched = true;
    s->c_loops++;
    return GVS_OK;
}

extern "C" int gvs_server_open(gvs_index *ix, const gvs_server_config *cfg, gvs_server **out)
{
    if (!ix || !cfg || !out || cfg->max_inflight < 1 || cfg->max_inflight > (1 << 20) || cfg->max_k < 1 ||
        cfg->max_ef < 1 || cfg->mc < 0 || cfg->storage != GVS_STORE_F32) {
        gvs::set_error("gvs_server_open: invalid configuration (fp32 storage only)");
        return GVS_ERR_ARG;
    }
    gvs_server *s = new gvs_server();
    s->ix = ix;
    s->cfg = *cfg;
    gvs_params p;
    std::memset(&p, 0, sizeof p);
    p.k = cfg->max_k;
    p.ef = cfg->max_ef;
    p.mc = cfg->mc > 0 ? cfg->mc : 1;
    p.mg = 1;
    p.waves_per_query = 1;
    p.storage = GVS_STORE_F32;
    p.persistent = 1;
    // the grid holds at most grid_waves (or max_inflight) waves: the
    // visited hash may use LDS that only waves beyond those would need
    int rc = gvs::plan_search(ix, &p, cfg->max_ef, cfg->max_k, &s->lp,
                              cfg->grid_waves > 0 ? cfg->grid_waves : cfg->max_inflight);
    if (rc) {
        delete s;
        return rc;
    }
    const int nch = ix->dim_pad / 32;
    const int rows = p.mc * ((ix->r0 + WAVE - 1) / WAVE);
    s->fn = ix->metric == GVS_METRIC_L2 ? gvs_serve_kernel_f32_l2(nch, rows) : gvs_serve_kernel_f32_ip(nch, rows);
    if (!s->fn) {
        delete s;
        gvs::set_error("gvs_server_open: dim %d not compiled for serving", ix->dim);
        return GVS_ERR_UNSUPPORTED;
    }
    int blocks_per_cu = 0;
    if (!gvs::hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, (const void *)s->fn,
                                                                  s->lp.wpb * WAVE,
                                                                  (size_t)s->lp.wpb * s->lp.lds_per_wave),
                     "occupancy")) {
        delete s;
        return GVS_ERR_HIP;
    }
    s->grid_waves = cfg->grid_waves > 0 ? cfg->grid_waves
                                        : (int64_t)std::max(1, blocks_per_cu) * ix->cu_count * s->lp.wpb;
    // a request occupies one wave, so waves beyond max_inflight could never
    // hold work; they would only add pollers
    if (cfg->grid_waves <= 0) s->grid_waves = std::min<int64_t>(s->grid_waves, (int64_t)cfg->max_inflight);
    s->grid_waves = (s->grid_waves + s->lp.wpb - 1) / s->lp.wpb * s->lp.wpb;
    const int64_t R = cfg->max_inflight;
    // ring of twice the slot count: an entry is reused only after the request
    // that held it is complete (checked at submit), so a holder never reads
    // an overwritten entry
    uint32_t rcap = 2;
    while (rcap < 2u * (uint32_t)R) rcap <<= 1;
    const int dim = ix->dim;
    bool ok = gvs::hip_ok(hipHostMalloc((void **)&s->ctl, sizeof(ServeCtl), hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->ring, 4 * (size_t)rcap, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qbuf, 4 * (size_t)R * dim, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qk, 4 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qef, 4 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->qseq, 8 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->done, 8 * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->cring, 8 * (size_t)rcap, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->labels, 8 * (size_t)R * cfg->max_k, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->dists, 4 * (size_t)R * cfg->max_k, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipHostMalloc((void **)&s->stats, 4 * GVS_STATS_ROW * (size_t)R, hipHostMallocCoherent), "hipHostMalloc") &&
              gvs::hip_ok(hipMalloc(&s->d_head, 8), "hipMalloc") && gvs::hip_ok(hipMalloc(&s->d_served, 8), "hipMalloc") &&
              gvs::hip_ok(hipMemset(s->d_head, 0, 8), "hipMemset") && gvs::hip_ok(hipMemset(s->d_served, 0, 8), "hipMemset") &&
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
    std::memset(s->cring, 0, 8 * (size_t)rcap);
    s->inflight_pos.assign((size_t)R, -1);
    std::memset(s->qseq, 0, 8 * (size_t)R);
    s->slots.assign((size_t)R, Slot());
    s->ring_busy.assign(rcap, 0);
    for (int64_t i = R - 1; i >= 0; i--) s->free_slots.push_back((int)i);
    KParams P;
    gvs::fill_params(ix, &p, s->lp, P);
    P.k = cfg->max_k;
    P.ef = cfg->max_ef;
    P.ldq = dim;
    P.out_labels = s->labels;
    P.out_dists = s->dists;
    P.out_stats = s->stats;
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
    s->args.ccap = rcap;
    s->args.head = s->d_head;
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

static int64_t submit_locked(gvs_server *s, const float *query, int32_t k, int32_t ef, double t_arrival)
{
    if (k < 1 || k > s->cfg.max_k || ef < 1 || ef > s->cfg.max_ef) {
        s->c_rejected++;
        gvs::set_error("gvs_server_submit: k=%d ef=%d outside the admitted range (max_k %d, max_ef %d)", k, ef,
                       s->cfg.max_k, s->cfg.max_ef);
        return GVS_ERR_ARG;
    }
    if (s->free_slots.empty() || s->ring_busy[s->tail & (uint64_t)(s->args.rcap - 1u)]) {
        gvs::set_error("gvs_server_submit: all %d device slots busy", s->cfg.max_inflight);
        return GVS_ERR_STATE;
    }
    const int slot = s->free_slots.back();
    s->free_slots.pop_back();
    std::memcpy(s->qbuf + (size_t)slot * s->ix->dim, query, sizeof(float) * s->ix->dim);
    s->qk[slot] = k;
    s->qef[slot] = ef;
    const uint64_t seq = s->next_seq++;
    s->qseq[slot] = seq;
    Slot &sl = s->slots[(size_t)slot];
    sl.req = s->next_req++;
    sl.t_arrival = t_arrival;
    sl.busy = true;
    sl.cancelled = false;
    s->inflight_pos[(size_t)slot] = (int)s->inflight.size();
    s->inflight.push_back(slot);
    sl.ring_idx = s->tail;
    s->ring_busy[s->tail & (uint64_t)(s->args.rcap - 1u)] = 1;
    s->ring[s->tail & (uint64_t)(s->args.rcap - 1u)] = (uint32_t)slot;
    s->tail++;
    // request fields and the ring entry are globally visible before the tail
    std::atomic_thread_fence(std::memory_order_seq_cst);
    __atomic_store_n(&s->ctl->tail, s->tail, __ATOMIC_RELEASE);
    s->c_submitted++;
    int rc = serve_ensure_running(s);
    if (rc) return rc;
    return sl.req;
}

extern "C" int64_t gvs_server_submit(gvs_server *s, const float *query, int32_t k, int32_t ef)
{
    if (!s || !query) {
        gvs::set_error("gvs_server_submit: invalid argument");
        return GVS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    return submit_locked(s, query, k, ef, now_us());
}

static int poll_locked(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels, float *dists,
                       double *latency_us, int32_t *stats_out)
{
    const int K = s->cfg.max_k;
    const uint64_t cmask = (uint64_t)s->args.ccap - 1u;
    int n = 0;
    // completions in the order the device finished them: entry i carries
    // (i + 1) in its high word once it is written
    while (n < max_done) {
        const uint64_t v = __atomic_load_n(&s->cring[s->chead & cmask], __ATOMIC_ACQUIRE);
        if ((uint32_t)(v >> 32) != (uint32_t)(s->chead + 1u)) break;
        const int slot = (int)(uint32_t)v;
        if (slot < 0 || slot >= s->cfg.max_inflight || !s->slots[(size_t)slot].busy) {
            gvs::set_error("gvs_server_poll: completion ring names slot %d, which is not in flight", slot);
            return GVS_ERR_STATE;
        }
        s->chead++;
        const double t = now_us();
        Slot &sl = s->slots[(size_t)slot];
        if (!sl.cancelled) {
            if (req_ids) req_ids[n] = sl.req;
            if (labels) std::memcpy(labels + (size_t)n * K, s->labels + (size_t)slot * K, 8 * (size_t)K);
            if (dists) std::memcpy(dists + (size_t)n * K, s->dists + (size_t)slot * K, 4 * (size_t)K);
            if (latency_us) latency_us[n] = t - sl.t_arrival;
            if (stats_out)
                std::memcpy(stats_out + (size_t)n * GVS_STATS_ROW, s->stats + (size_t)slot * GVS_STATS_ROW,
                            4 * GVS_STATS_ROW);
            n++;
            s->c_completed++;
        }
        sl.busy = false;
        s->ring_busy[sl.ring_idx & (uint64_t)(s->args.rcap - 1u)] = 0;
        s->free_slots.push_back(slot);
        const int ip = s->inflight_pos[(size_t)slot];
        const int last = s->inflight.back();
        s->inflight[(size_t)ip] = last;
        s->inflight_pos[(size_t)last] = ip;
        s->inflight.pop_back();
        s->inflight_pos[(size_t)slot] = -1;
    }
    if (!s->inflight.empty()) {
        int rc = serve_ensure_running(s);  // a launch that went idle leaves work behind
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
    std::lock_guard<std::mutex> lk(s->mu);
    return poll_locked(s, max_done, req_ids, labels, dists, latency_us, nullptr);
}

extern "C" int gvs_server_poll_stats(gvs_server *s, int32_t max_done, int64_t *req_ids, int64_t *labels,
                                     float *dists, double *latency_us, int32_t *stats)
{
    if (!s || max_done < 0) {
        gvs::set_error("gvs_server_poll_stats: invalid argument");
        return GVS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    return poll_locked(s, max_done, req_ids, labels, dists, latency_us, stats);
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

extern "C" int gvs_server_close(gvs_server *s)
{
    if (!s) return GVS_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        __atomic_store_n(&s->ctl->stop, 1u, __ATOMIC_RELEASE);
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
// (rate_qps) and is submitted as soon as it has arrived and a slot is free;
// its latency runs from its arrival to the poll that sees it completed.
extern "C" int gvs_server_loadgen(gvs_server *s, const float *queries, int64_t nq, int64_t ldq, double rate_qps,
                                  uint64_t seed, const int32_t *ks, const int32_t *efs, int64_t *out_labels,
                                  float *out_dists, double *out_latency_us, double *out_wall_s)
{
    if (!s || !queries || nq < 0 || ldq < s->ix->dim || rate_qps <= 0.0 || !ks || !efs) {
        gvs::set_error("gvs_server_loadgen: invalid argument");
        return GVS_ERR_ARG;
    }
    const int K = s->cfg.max_k;
    std::vector<double> arrival((size_t)nq);
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_qps / 1e6);  // per microsecond
    double t = 0.0;
    for (int64_t i = 0; i < nq; i++) {
        t += gap(rng);
        arrival[(size_t)i] = t;
    }
    std::vector<int64_t> req_of((size_t)nq, -1);
    std::vector<int64_t> index_of_req;
    index_of_req.reserve((size_t)nq);
    const int batch = 256;
    std::vector<int64_t> rid(batch), lab((size_t)batch * K);
    std::vector<float> dd((size_t)batch * K);
    std::vector<double> lat(batch);
    const double t0 = now_us();
    int64_t next = 0, done_n = 0;
    const int64_t req_base = s->next_req;
    double last_progress = 0.0;
    while (done_n < nq) {
        const double now = now_us() - t0;
        if (now - last_progress > 30e6) {
            gvs::set_error("gvs_server_loadgen: no request completed for 30 s (%lld of %lld done)",
                           (long long)done_n, (long long)nq);
            return GVS_ERR_STATE;
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (next < nq && arrival[(size_t)next] <= now) {
                int64_t r = submit_locked(s, queries + next * ldq, ks[next], efs[next], t0 + arrival[(size_t)next]);
                if (r == GVS_ERR_STATE) break;  // no free slot: retry after completions
                if (r < 0) return (int)r;
                req_of[(size_t)next] = r;
                next++;
            }
            int n = poll_locked(s, batch, rid.data(), lab.data(), dd.data(), lat.data(), nullptr);
            if (n < 0) return n;
            for (int j = 0; j < n; j++) {
                const int64_t i = rid[(size_t)j] - req_base;  // requests are numbered in submission order
                if (i < 0 || i >= nq) continue;
                if (out_labels) std::memcpy(out_labels + i * K, lab.data() + (size_t)j * K, 8 * (size_t)K);
                if (out_dists) std::memcpy(out_dists + i * K, dd.data() + (size_t)j * K, 4 * (size_t)K);
                if (out_latency_us) out_latency_us[i] = lat[(size_t)j];
                done_n++;
            }
            if (n > 0) last_progress = now;
        }
    }
    if (out_wall_s) *out_wall_s = (now_us() - t0) / 1e6;
    return GVS_OK;
}
