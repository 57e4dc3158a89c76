// Device index, search launch and the C ABI entry points for search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "gvs_internal.hpp"
#include "search_kernels.hpp"

using namespace gvsk;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// work-queue head and overflow counter reset (a kernel, not a memset, so the
// launch sequence can be captured in a hipGraph)
__global__ void k_reset_counters(uint32_t *c)
{
    if (threadIdx.x < 12) c[threadIdx.x] = 0u;
}

// the Bloom filter's hash of n keys (known-answer checks)
__global__ __launch_bounds__(256) void k_murmur2_k4(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ seeds,
                                                    uint32_t *__restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = murmur2_k4(keys[i], seeds[i]);
}

// plain streaming copy (16 bytes per lane per iteration)
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// seq: fp32 rows summed in hnswlib's scalar order (GVS_FLAG_HNSWLIB_ORDER)
// Device memory of the no-allocation search (gvs_search_device_noalloc):
// the per-query status words, then the per-wave visited tables.  One copy
// per GPU, held until the process ends; host code never passes `scratch`
// itself to a kernel, only the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch below).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

__global__ __launch_bounds__(256) void k_zero16(uint4 *__restrict__ dst, int64_t n16)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = z;
}

static dfn_t pick_descend(int storage, int metric, int nch, bool seq)
{
    const bool l2 = metric == GVS_METRIC_L2;
    if (seq) return storage == GVS_STORE_F32 ? (l2 ? gvs_descend_f32seq_l2(nch) : gvs_descend_f32seq_ip(nch)) : nullptr;
    if (storage == GVS_STORE_F32) return l2 ? gvs_descend_f32_l2(nch) : gvs_descend_f32_ip(nch);
    if (storage == GVS_STORE_BF16) return l2 ? gvs_descend_bf16_l2(nch) : gvs_descend_bf16_ip(nch);
    if (storage == GVS_STORE_FP8) return l2 ? gvs_descend_fp8_l2(nch) : gvs_descend_fp8_ip(nch);
    return nullptr;
}

static kfn_t pick_kernel(int storage, int metric, int nch, int rows, bool seq, bool pool)
{
    const bool l2 = metric == GVS_METRIC_L2;
    if (pool) {
        if (seq) return nullptr;
        if (storage == GVS_STORE_F32) return l2 ? gvs_kernel_pool_f32_l2(nch, rows) : gvs_kernel_pool_f32_ip(nch, rows);
        if (storage == GVS_STORE_BF16) return l2 ? gvs_kernel_pool_bf16_l2(nch, rows) : gvs_kernel_pool_bf16_ip(nch, rows);
        if (storage == GVS_STORE_FP8) return l2 ? gvs_kernel_pool_fp8_l2(nch, rows) : gvs_kernel_pool_fp8_ip(nch, rows);
        return nullptr;
    }
    if (seq) return storage == GVS_STORE_F32 ? (l2 ? gvs_kernel_f32seq_l2(nch, rows) : gvs_kernel_f32seq_ip(nch, rows)) : nullptr;
    if (storage == GVS_STORE_F32) return l2 ? gvs_kernel_f32_l2(nch, rows) : gvs_kernel_f32_ip(nch, rows);
    if (storage == GVS_STORE_BF16) return l2 ? gvs_kernel_bf16_l2(nch, rows) : gvs_kernel_bf16_ip(nch, rows);
    if (storage == GVS_STORE_FP8) return l2 ? gvs_kernel_fp8_l2(nch, rows) : gvs_kernel_fp8_ip(nch, rows);
    return nullptr;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
namespace gvs {

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_error("%s: %s", what, hipGetErrorString(e));
    return false;
}

static int ilog2_ceil(int64_t v)
{
    int l = 0;
    while ((int64_t(1) << l) < v) l++;
    return l;
}

// bf16 / e4m3 encoders for the low-precision copies (round to nearest even)
static uint16_t f32_to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float e4m3_value(int c)
{
    int e = (c >> 3) & 15, m = c & 7;
    return e ? std::ldexp(1.0f + m / 8.0f, e - 7) : (float)m * 0x1p-9f;
}
static uint8_t f32_to_e4m3(float f)
{
    uint8_t s = std::signbit(f) ? 0x80 : 0;
    float a = std::fabs(f);
    if (a >= 448.0f) return s | 0x7E;
    // binary search over the monotone grid of codes 0..0x7E
    int lo = 0, hi = 0x7E;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (e4m3_value(mid) <= a) lo = mid;
        else hi = mid;
    }
    float dl = a - e4m3_value(lo), dh = e4m3_value(hi) - a;
    int c = (dl < dh) ? lo : (dh < dl) ? hi : ((lo & 1) ? hi : lo);
    return s | (uint8_t)c;
}

}  // namespace gvs

using gvs::hip_ok;

namespace gvs {
// Address of this device's `scratch` array, resolved once per device with
// that device current (gvs_index_create is the set-up call).
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[64] = {};

static int noalloc_scratch(int device, unsigned char **out)
{
    if (device < 0 || device >= 64) {
        set_error("gvs: device %d out of range for the no-allocation scratch", device);
        return GVS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !ptr)
            return GVS_ERR_HIP;
        g_scratch_dev[device] = (unsigned char *)ptr;
    }
    *out = g_scratch_dev[device];
    return GVS_OK;
}
}  // namespace gvs

static void free_index(gvs_index *ix)
{
    if (!ix) return;
    if (!ix->parent) {  // a fork's arrays are its root's
        (void)hipFree(ix->d_f32);
        (void)hipFree(ix->d_bf16);
        (void)hipFree(ix->d_fp8);
        (void)hipFree(ix->d_scale);
        (void)hipFree(ix->d_adj0);
        (void)hipFree(ix->d_adj_up);
        (void)hipFree(ix->d_upper_off);
        (void)hipFree(ix->d_labels);
    }
    (void)hipFree(ix->d_vtab);
    (void)hipFree(ix->d_counters);
    (void)hipFree(ix->d_status);
    (void)hipFree(ix->d_xnorm64);
    (void)hipFree(ix->d_rep);
    (void)hipFree(ix->d_rkey);
    (void)hipFree(ix->d_rperm);
    (void)hipFree(ix->d_rtmp);
    delete ix;
}

extern "C" int gvs_index_create(const gvs_graph *g, uint32_t storage_mask, gvs_index **out)
{
    if (!g || !out || g->n <= 0 || g->n > INT32_MAX || g->dim <= 0 || !g->vectors || !g->adj0 ||
        g->r0 <= 0 || g->ld_vectors < g->dim || g->entry < 0 || g->entry >= g->n ||
        (g->metric != GVS_METRIC_L2 && g->metric != GVS_METRIC_IP) || g->max_level < 0) {
        gvs::set_error("gvs_index_create: invalid graph description");
        return GVS_ERR_ARG;
    }
    if (g->r0 > WAVE * MAX_ROWS) {
        gvs::set_error("gvs_index_create: r0=%d exceeds %d", g->r0, WAVE * MAX_ROWS);
        return GVS_ERR_UNSUPPORTED;
    }
    if (g->max_level > 0 && (!g->upper_off || !g->adj_up || g->ru <= 0 || g->ru > WAVE)) {
        gvs::set_error("gvs_index_create: upper layers need upper_off, adj_up and 0 < ru <= 64");
        return GVS_ERR_ARG;
    }
    // validate ids (a bad id would make the kernel read out of bounds)
    const int64_t n = g->n;
    for (int64_t i = 0; i < n * g->r0; i++)
        if (g->adj0[i] < -1 || g->adj0[i] >= n) {
            gvs::set_error("gvs_index_create: adj0[%lld] = %d out of range", (long long)i, g->adj0[i]);
            return GVS_ERR_ARG;
        }
    if (g->max_level > 0) {
        for (int64_t i = 0; i < g->up_rows * g->ru; i++)
            if (g->adj_up[i] < -1 || g->adj_up[i] >= n) {
                gvs::set_error("gvs_index_create: adj_up[%lld] out of range", (long long)i);
                return GVS_ERR_ARG;
            }
        for (int64_t i = 0; i < n; i++)
            if (g->upper_off[i] < -1 || g->upper_off[i] >= g->up_rows) {
                gvs::set_error("gvs_index_create: upper_off[%lld] out of range", (long long)i);
                return GVS_ERR_ARG;
            }
        // node levels from the row layout (rows of a node are consecutive);
        // the descent at level l only follows links to nodes of level >= l
        std::vector<std::pair<int64_t, int64_t>> offs;
        for (int64_t i = 0; i < n; i++)
            if (g->upper_off[i] >= 0) offs.emplace_back(g->upper_off[i], i);
        std::sort(offs.begin(), offs.end());
        std::vector<int32_t> lev((size_t)n, 0);
        std::vector<int64_t> owner((size_t)g->up_rows, -1);
        for (size_t j = 0; j < offs.size(); j++) {
            int64_t end = j + 1 < offs.size() ? offs[j + 1].first : g->up_rows;
            lev[(size_t)offs[j].second] = (int32_t)(end - offs[j].first);
            for (int64_t r = offs[j].first; r < end; r++) owner[(size_t)r] = offs[j].second;
        }
        if (lev[(size_t)g->entry] < g->max_level) {
            gvs::set_error("gvs_index_create: entry point level %d < max_level %d", lev[(size_t)g->entry],
                           g->max_level);
            return GVS_ERR_ARG;
        }
        for (int64_t r = 0; r < g->up_rows; r++) {
            if (owner[(size_t)r] < 0) continue;
            int l = (int)(r - g->upper_off[(size_t)owner[(size_t)r]]) + 1;
            for (int j = 0; j < g->ru; j++) {
                int32_t u = g->adj_up[r * g->ru + j];
                if (u >= 0 && lev[(size_t)u] < l) {
                    gvs::set_error("gvs_index_create: level-%d link to node %d of level %d", l, u, lev[(size_t)u]);
                    return GVS_ERR_ARG;
                }
            }
        }
    }

    gvs_index *ix = new gvs_index();
    ix->n = n;
    ix->dim = g->dim;
    ix->metric = g->metric;
    ix->r0 = g->r0;
    ix->ru = g->ru;
    ix->max_level = g->max_level;
    ix->entry = g->entry;
    ix->up_rows = g->max_level > 0 ? g->up_rows : 0;
    ix->storage_mask = storage_mask | (1u << GVS_STORE_F32);
    int gran = (ix->storage_mask & (1u << GVS_STORE_FP8)) ? 128 : (ix->storage_mask & (1u << GVS_STORE_BF16)) ? 64 : 32;
    ix->dim_pad = (g->dim + gran - 1) / gran * gran;
    (void)hipGetDevice(&ix->device);
    {
        unsigned char *scr = nullptr;  // set-up of the no-allocation search on this device
        int rc = gvs::noalloc_scratch(ix->device, &scr);
        if (rc) {
            delete ix;
            return rc;
        }
    }
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, ix->device), "hipGetDeviceProperties")) {
        delete ix;
        return GVS_ERR_HIP;
    }
    ix->cu_count = prop.multiProcessorCount;

    const int dp = ix->dim_pad;
    std::vector<float> padded;
    bool ok = true;
    // fp32 copy, zero padded rows
    {
        padded.assign((size_t)n * dp, 0.0f);
        for (int64_t i = 0; i < n; i++)
            std::memcpy(&padded[(size_t)i * dp], g->vectors + i * g->ld_vectors, sizeof(float) * g->dim);
        ok = ok && hip_ok(hipMalloc(&ix->d_f32, padded.size() * 4), "hipMalloc vectors") &&
             hip_ok(hipMemcpy(ix->d_f32, padded.data(), padded.size() * 4, hipMemcpyHostToDevice), "copy vectors");
    }
    if (ok && (ix->storage_mask & (1u << GVS_STORE_BF16))) {
        std::vector<uint16_t> h((size_t)n * dp);
        for (size_t i = 0; i < h.size(); i++) h[i] = gvs::f32_to_bf16(padded[i]);
        ok = hip_ok(hipMalloc(&ix->d_bf16, h.size() * 2), "hipMalloc bf16") &&
             hip_ok(hipMemcpy(ix->d_bf16, h.data(), h.size() * 2, hipMemcpyHostToDevice), "copy bf16");
    }
    if (ok && (ix->storage_mask & (1u << GVS_STORE_FP8))) {
        std::vector<uint8_t> h((size_t)n * dp);
        std::vector<float> sc((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            const float *row = &padded[(size_t)i * dp];
            float amax = 0.0f;
            for (int d = 0; d < dp; d++) amax = std::max(amax, std::fabs(row[d]));
            float s = amax > 0.0f ? amax / 448.0f : 1.0f;
            sc[(size_t)i] = s;
            for (int d = 0; d < dp; d++) h[(size_t)i * dp + d] = gvs::f32_to_e4m3(row[d] / s);
        }
        ok = hip_ok(hipMalloc(&ix->d_fp8, h.size()), "hipMalloc fp8") &&
             hip_ok(hipMemcpy(ix->d_fp8, h.data(), h.size(), hipMemcpyHostToDevice), "copy fp8") &&
             hip_ok(hipMalloc(&ix->d_scale, sc.size() * 4), "hipMalloc scale") &&
             hip_ok(hipMemcpy(ix->d_scale, sc.data(), sc.size() * 4, hipMemcpyHostToDevice), "copy scale");
    }
    padded.clear();
    padded.shrink_to_fit();
    ok = ok && hip_ok(hipMalloc(&ix->d_adj0, sizeof(int32_t) * n * g->r0), "hipMalloc adj0") &&
         hip_ok(hipMemcpy(ix->d_adj0, g->adj0, sizeof(int32_t) * n * g->r0, hipMemcpyHostToDevice), "copy adj0");
    if (ok && g->max_level > 0) {
        ok = hip_ok(hipMalloc(&ix->d_adj_up, sizeof(int32_t) * g->up_rows * g->ru), "hipMalloc adj_up") &&
             hip_ok(hipMemcpy(ix->d_adj_up, g->adj_up, sizeof(int32_t) * g->up_rows * g->ru, hipMemcpyHostToDevice),
                    "copy adj_up") &&
             hip_ok(hipMalloc(&ix->d_upper_off, sizeof(int64_t) * n), "hipMalloc upper_off") &&
             hip_ok(hipMemcpy(ix->d_upper_off, g->upper_off, sizeof(int64_t) * n, hipMemcpyHostToDevice),
                    "copy upper_off");
    }
    if (ok && g->labels) {
        ok = hip_ok(hipMalloc(&ix->d_labels, sizeof(int64_t) * n), "hipMalloc labels") &&
             hip_ok(hipMemcpy(ix->d_labels, g->labels, sizeof(int64_t) * n, hipMemcpyHostToDevice), "copy labels");
    }
    ok = ok && hip_ok(hipMalloc(&ix->d_counters, 64), "hipMalloc counters") &&
         hip_ok(hipMemset(ix->d_counters, 0, 64), "memset counters");
    if (!ok) {
        free_index(ix);
        return GVS_ERR_HIP;
    }
    *out = ix;
    return GVS_OK;
}

// guards the fork counts of root indexes
static std::mutex g_fork_mu;

extern "C" int gvs_index_destroy(gvs_index *ix)
{
    if (!ix) return GVS_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(g_fork_mu);
        if (ix->forks > 0) {
            gvs::set_error("gvs_index_destroy: %d forks of this index are still open", ix->forks);
            return GVS_ERR_STATE;
        }
        if (ix->parent) ix->parent->forks--;
    }
    (void)hipDeviceSynchronize();
    free_index(ix);
    return GVS_OK;
}

extern "C" int gvs_index_fork(gvs_index *ix, gvs_index **out)
{
    if (!ix || !out) {
        gvs::set_error("gvs_index_fork: invalid argument");
        return GVS_ERR_ARG;
    }
    gvs_index *root = ix->parent ? ix->parent : ix;
    int dev = -1;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return GVS_ERR_HIP;
    if (dev != root->device) {
        gvs::set_error("gvs_index_fork: the index lives on device %d, the current device is %d", root->device, dev);
        return GVS_ERR_ARG;
    }
    gvs_index *f = new gvs_index(*root);
    // own scratch only (allocated on demand, as for a new index)
    f->parent = root;
    f->forks = 0;
    f->d_vtab = nullptr;
    f->vtab_words = 0;
    f->d_counters = nullptr;
    f->d_status = nullptr;
    f->status_cap = 0;
    f->d_xnorm64 = nullptr;
    f->d_rep = nullptr;
    f->d_rkey = nullptr;
    f->d_rperm = nullptr;
    f->d_rtmp = nullptr;
    f->rcap = 0;
    f->rtmp_bytes = 0;
    if (!hip_ok(hipMalloc(&f->d_counters, 64), "hipMalloc counters") ||
        !hip_ok(hipMemset(f->d_counters, 0, 64), "memset counters")) {
        (void)hipFree(f->d_counters);
        delete f;
        return GVS_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(g_fork_mu);
        root->forks++;
    }
    *out = f;
    return GVS_OK;
}

extern "C" int gvs_index_device_bytes(const gvs_index *ix, int64_t *graph_bytes, int64_t *vector_bytes)
{
    if (!ix) return GVS_ERR_ARG;
    if (graph_bytes)
        *graph_bytes = ix->n * ix->r0 * 4 + ix->up_rows * ix->ru * 4 + (ix->max_level > 0 ? ix->n * 8 : 0);
    if (vector_bytes) {
        int64_t v = ix->n * ix->dim_pad * 4;
        if (ix->d_bf16) v += ix->n * ix->dim_pad * 2;
        if (ix->d_fp8) v += ix->n * ix->dim_pad + ix->n * 4;
        *vector_bytes = v;
    }
    return GVS_OK;
}

// ---------------------------------------------------------------------------
// launch planning
// ---------------------------------------------------------------------------
namespace gvs {

// One wave per query: LDS per wave, waves per block (the most resident waves
// per CU: 160 KiB LDS, at most 16 workgroups per CU, at most 64 KiB per
// block) and the resident wave count, for lp's tables.
static int fit_waves(const gvs_index *ix, const void *fn, LaunchPlan *lp)
{
    int bytes = 16 * lp->ef_cap + 8 * lp->lcap + 8 * (1 << lp->blog2) + 4 * lp->lcap + 4 * gvsk::gvs_sel_words(lp->mg) +
                (lp->hlog2 > 0 ? (lp->hnarrow ? 2 : 4) * (1 << lp->hlog2) : 0);
    lp->lds_per_wave = (bytes + 15) / 16 * 16;
    lp->wpb = 1;
    int best = 0;
    for (int w = 1; w <= 4; w *= 2) {
        if (w * lp->lds_per_wave > 65536) break;
        int per_cu = std::min(16, (160 * 1024) / (w * lp->lds_per_wave)) * w;
        if (per_cu >= best) {
            best = per_cu;
            lp->wpb = w;
        }
    }
    if (lp->lds_per_wave > 65536) {
        set_error("gvs_search: per-query LDS %d bytes exceeds 64 KiB (ef too large)", lp->lds_per_wave);
        return GVS_ERR_UNSUPPORTED;
    }
    int blocks_per_cu = 0;
    if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, fn, lp->wpb * WAVE,
                                                             (size_t)lp->wpb * lp->lds_per_wave),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return GVS_ERR_HIP;
    lp->resident_waves = std::max(1, blocks_per_cu) * ix->cu_count * lp->wpb;
    return GVS_OK;
}

// max_waves > 0: the launch runs at most that many waves (a server's grid),
// so resident waves beyond it do not count when the visited hash is sized.
int plan_search(const gvs_index *ix, const gvs_params *p, int max_ef, int max_k, LaunchPlan *lp,
                int64_t max_waves = 0, bool pool = false)
{
    if (!p) {
        set_error("gvs_search: params is NULL");
        return GVS_ERR_ARG;
    }
    if (max_k < 1 || max_k > 4096 || max_ef < 1 || max_ef > 4096 || p->mc < 1 || p->mc > 8 || p->mg < 1 ||
        p->mg > MAX_MG || (p->waves_per_query != 1 && p->waves_per_query != 2 && p->waves_per_query != 4 &&
                      p->waves_per_query != 8)) {
        set_error("gvs_search: unsupported k=%d ef=%d mc=%d mg=%d waves_per_query=%d", max_k, max_ef, p->mc,
                  p->mg, p->waves_per_query);
        return GVS_ERR_UNSUPPORTED;
    }
    if (p->storage < 0 || p->storage > 2 || !(ix->storage_mask & (1u << p->storage))) {
        set_error("gvs_search: storage %d not built for this index", p->storage);
        return GVS_ERR_ARG;
    }
    const int rpc = (ix->r0 + WAVE - 1) / WAVE;
    if (p->mc * rpc > MAX_ROWS) {
        set_error("gvs_search: mc * ceil(r0/64) must be <= %d", MAX_ROWS);
        return GVS_ERR_UNSUPPORTED;
    }
    const int E = p->storage == GVS_STORE_F32 ? 4 : p->storage == GVS_STORE_BF16 ? 8 : 16;
    const int nch = ix->dim_pad / (8 * E);
    const int W = p->waves_per_query;
    int wrows = 0;  // adjacency rows per wave when several waves share a query
    if (W > 1) {
        if (p->storage != GVS_STORE_F32) {
            set_error("gvs_search: waves_per_query > 1 needs fp32 storage");
            return GVS_ERR_UNSUPPORTED;
        }
        const int rpw = (p->mc * rpc + W - 1) / W;
        wrows = rpw <= 1 ? 1 : rpw <= 2 ? 2 : 4;
    }
    const bool seq = (p->flags & GVS_FLAG_HNSWLIB_ORDER) != 0;
    if (seq && p->storage != GVS_STORE_F32) {
        set_error("gvs_search: GVS_FLAG_HNSWLIB_ORDER needs fp32 storage");
        return GVS_ERR_UNSUPPORTED;
    }
    const bool l2 = ix->metric == GVS_METRIC_L2;
    kfn_t fn = nullptr;
    if (ix->dim_pad % (8 * E) == 0)
        fn = W > 1 ? (seq ? (l2 ? gvs_kernel_mw_f32seq_l2(nch, wrows) : gvs_kernel_mw_f32seq_ip(nch, wrows))
                          : (l2 ? gvs_kernel_mw_f32_l2(nch, wrows) : gvs_kernel_mw_f32_ip(nch, wrows)))
                   : pick_kernel(p->storage, ix->metric, nch, p->mc * rpc, seq, pool);
    if (!fn) {
        set_error("gvs_search: dim %d (padded %d) not compiled for storage %d", ix->dim, ix->dim_pad, p->storage);
        return GVS_ERR_UNSUPPORTED;
    }
    int efe = std::max(max_ef, max_k);
    lp->fn = (void *)fn;
    lp->dfn = (void *)pick_descend(p->storage, ix->metric, nch, seq);
    lp->mg = p->mg;
    lp->ef_cap = (efe + 63) / 64 * 64;
    lp->lcap = p->mc * rpc * WAVE;
    lp->blog2 = p->bloom_log2 > 0 ? p->bloom_log2 : std::min(12, std::max(8, ilog2_ceil((int64_t)efe * 8)));
    int vmin = ilog2_ceil(4 * (int64_t)lp->lcap + 1024);
    // with GVS_FLAG_LDS_VISITED, visited_log2 in [10, 14] sizes the LDS hash
    // and the HBM table (used once that hash is frozen) keeps its default
    const bool lds_sized = (p->flags & GVS_FLAG_LDS_VISITED) && p->visited_log2 >= 10 && p->visited_log2 <= 14;
    lp->vlog2 = p->visited_log2 > 0 && !lds_sized ? std::max(p->visited_log2, vmin)
                                                   : std::max(vmin, std::min(20, ilog2_ceil((int64_t)efe * 128)));
    // exact visited set in LDS (no HBM table until it fills), unless disabled
    lp->hlog2 = 0;
    // ef in (128, 512]: an exact LDS hash of 2-byte slots (a node's quotient
    // and probe step), when a node's quotient fits 8 bits: 2^14 slots, or
    // 2^13 where 2^14 would cost waves the launch can use, or 2^13 4-byte
    // slots in the same 32 KiB where those cost none (below).  Few queries
    // fill 2^13 slots to 3/4 at ef 512 on degree-64 graphs; those go on
    // with the slower frozen lookups plus the HBM table.
    const int nlog2 = lds_sized ? p->visited_log2 : 14;
    const int nbits = std::max(nlog2, ilog2_ceil(std::max<int64_t>(ix->n, 2)));
    const bool wide = (p->flags & GVS_FLAG_LDS_VISITED) && (p->flags & GVS_FLAG_LDS_WIDE);
    const bool narrow = !(p->flags & GVS_FLAG_HBM_VISITED) && !wide && W == 1 && efe > 128 && nbits <= nlog2 + 8 &&
                        ((p->flags & GVS_FLAG_LDS_VISITED) || efe <= 512);
    if (narrow) {
        lp->hlog2 = nlog2;
        lp->hnarrow = 1;
        lp->hbits = nbits;
        lp->blog2 = p->bloom_log2 > 0 ? p->bloom_log2 : 1;
    } else if (!(p->flags & GVS_FLAG_HBM_VISITED) && W > 1 && efe > 128 && efe <= 512 && !lds_sized) {
        // several waves per query share one workgroup's LDS (up to 160 KiB):
        // an exact 4-byte hash of 2^13 slots (2^14 above ef 256), which holds
        // the evaluations of such queries, frozen and spilled to the HBM
        // table as the one-wave hashes are
        lp->hlog2 = efe > 256 ? 14 : 13;
        lp->blog2 = p->bloom_log2 > 0 ? p->bloom_log2 : 1;
    } else if (!(p->flags & GVS_FLAG_HBM_VISITED) && ((p->flags & GVS_FLAG_LDS_VISITED) || efe <= 128)) {
        // 4-byte slots (2^12 hold the p99 query at ef 64-96 on degree-64
        // graphs; in a full batch 2^13 costs more in occupancy than the
        // spills it saves)
        int want = ilog2_ceil(((int64_t)efe * 40 + lp->lcap) * 4 / 3);
        lp->hlog2 = lds_sized ? p->visited_log2 : std::max(11, std::min(12, want));
        // Bloom filter for the ids met after the hash is frozen; by default
        // none (2 words keep the hash 16-byte aligned): those ids are then
        // all gathered speculatively
        lp->blog2 = p->bloom_log2 > 0 ? p->bloom_log2 : 1;
    }
    if (lp->blog2 > 13 || lp->vlog2 > 26) {
        set_error("gvs_search: bloom_log2 <= 13 and visited_log2 <= 26 required");
        return GVS_ERR_ARG;
    }
    if (W > 1) {
        // one workgroup per query: shared pool, Bloom words, selection and
        // control words, visited hash; per wave wrows*64 keys and ids
        int bytes = 16 * lp->ef_cap + 8 * (1 << lp->blog2) + W * wrows * WAVE * 12 + 32 * p->mg + 4 * WG_CTL_WORDS +
                    (lp->hlog2 > 0 ? 4 * (1 << lp->hlog2) : 0);
        lp->lds_per_wave = (bytes + 15) / 16 * 16;  // per workgroup
        lp->wpb = W;
        if (lp->lds_per_wave > 160 * 1024 - 64) {  // (the persistent loop's claim word is static LDS)
            set_error("gvs_search: per-query LDS %d bytes exceeds 160 KiB (ef too large)", lp->lds_per_wave);
            return GVS_ERR_UNSUPPORTED;
        }
        if (lp->lds_per_wave > 65536 &&
            !hip_ok(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        lp->lds_per_wave),
                    "hipFuncSetAttribute"))
            return GVS_ERR_HIP;
        // resident workgroups (the persistent grid), in waves
        int bpc = 0;
        if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void *)fn, W * WAVE, lp->lds_per_wave),
                    "occupancy"))
            return GVS_ERR_HIP;
        lp->resident_waves = std::max(1, bpc) * ix->cu_count * W;
        return GVS_OK;
    }
    const int rc = fit_waves(ix, (const void *)fn, lp);
    // A default-sized LDS hash takes the larger of two sizes unless that
    // costs waves the launch can use: the shorter probe sequences of an
    // emptier table help, but a chain-bound search gains more from waves.
    const auto waves = [&](const LaunchPlan &l) {
        return max_waves > 0 ? std::min<int64_t>(l.resident_waves, max_waves) : (int64_t)l.resident_waves;
    };
    if (narrow && !lds_sized && lp->hlog2 == 14) {
        // 2-byte slots: 2^14, else 2^13
        LaunchPlan best = *lp;
        int best_rc = rc;
        LaunchPlan alt = *lp;
        alt.hlog2 = 13;
        alt.hbits = std::max(13, ilog2_ceil(std::max<int64_t>(ix->n, 2)));
        if (alt.hbits <= 13 + 8 && fit_waves(ix, (const void *)fn, &alt) == GVS_OK &&
            (best_rc != GVS_OK || waves(alt) > waves(best))) {
            best = alt;
            best_rc = GVS_OK;
        }
        // 2^13 4-byte slots (32 KiB, as 2^14 2-byte ones) where they cost no
        // waves the launch can use: a 4-byte slot is checked with fewer
        // dependent instructions, which few waves feel most
        LaunchPlan a4 = *lp;
        a4.hlog2 = 13;
        a4.hnarrow = 0;
        a4.hbits = 0;
        if (best_rc == GVS_OK && fit_waves(ix, (const void *)fn, &a4) == GVS_OK && waves(a4) >= waves(best))
            best = a4;
        *lp = best;
        return best_rc;
    } else if (!narrow && !lds_sized && W == 1 && lp->hlog2 > 0 && lp->hlog2 < 13 && rc == GVS_OK) {
        // 4-byte slots: the size ef asks for, or twice that
        LaunchPlan alt = *lp;
        alt.hlog2 = lp->hlog2 + 1;
        if (fit_waves(ix, (const void *)fn, &alt) == GVS_OK && waves(alt) >= waves(*lp)) {
            *lp = alt;
            return GVS_OK;
        }
    }
    return rc;
}

int ensure_scratch(gvs_index *ix, int64_t slots, int vlog2, int64_t nstatus)
{
    size_t words = (size_t)slots << vlog2;
    if (words > ix->vtab_words) {
        (void)hipFree(ix->d_vtab);
        ix->d_vtab = nullptr;
        ix->vtab_words = 0;
        if (!hip_ok(hipMalloc(&ix->d_vtab, words * 4), "hipMalloc visited tables") ||
            !hip_ok(hipMemset(ix->d_vtab, 0, words * 4), "hipMemset visited tables"))
            return GVS_ERR_NOMEM;
        ix->vtab_words = words;
    }
    if (nstatus > ix->status_cap) {
        (void)hipFree(ix->d_status);
        ix->d_status = nullptr;
        ix->status_cap = 0;
        if (!hip_ok(hipMalloc(&ix->d_status, sizeof(int32_t) * nstatus), "hipMalloc status")) return GVS_ERR_NOMEM;
        ix->status_cap = nstatus;
    }
    return GVS_OK;
}

void fill_params(const gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, KParams &P)
{
    std::memset(&P, 0, sizeof P);
    P.vec = p->storage == GVS_STORE_F32 ? (const void *)ix->d_f32
          : p->storage == GVS_STORE_BF16 ? (const void *)ix->d_bf16 : (const void *)ix->d_fp8;
    P.vscale = ix->d_scale;
    P.vec_f32 = ix->d_f32;
    P.row_elems = ix->dim_pad;
    P.row_elems_f32 = ix->dim_pad;
    P.dim = ix->dim;
    P.r0 = ix->r0;
    P.adj0 = ix->d_adj0;
    P.adj_up = ix->d_adj_up;
    P.upper_off = ix->d_upper_off;
    P.ru = ix->ru;
    P.max_level = ix->max_level;
    P.entry = ix->entry;
    P.labels = ix->d_labels;
    P.n = ix->n;
    P.mc = p->mc;
    P.mg = p->mg;
    P.rerank = p->rerank;
    P.vtab = ix->d_vtab;
    P.vlog2 = lp.vlog2;
    P.blog2 = lp.blog2;
    P.hlog2 = lp.hlog2;
    P.hnarrow = lp.hnarrow;
    P.hbits = lp.hbits;
    P.bloom_seed = 0x2545F491u;
    P.ef_cap = lp.ef_cap;
    P.lcap = lp.lcap;
    P.lds_per_wave = lp.lds_per_wave;
    P.waves_per_block = lp.wpb;
    P.counters = ix->d_counters;
    P.status = ix->d_status;
}

// Waves that need a visited table for a search of nq queries: the resident
// grid in persistent mode, else the chunk size of the one-wave-per-query
// launches.
int64_t scratch_slots(const gvs_params *p, const LaunchPlan &lp, int64_t nq)
{
    if (p->waves_per_query > 1) {
        if (p->persistent) {  // the resident workgroups, one table each
            const int64_t gw = p->grid_waves > 0 ? std::min<int64_t>(p->grid_waves, lp.resident_waves) : lp.resident_waves;
            return std::max<int64_t>(1, std::min<int64_t>(nq, gw / lp.wpb));
        }
        // one workgroup (one table) per query, in chunks
        return std::min<int64_t>(std::min<int64_t>(nq, 65536), std::max<int64_t>(1, ((int64_t)1 << 30) >> lp.vlog2));
    }
    if (p->persistent) {
        int64_t gw = p->grid_waves > 0 ? p->grid_waves : lp.resident_waves;
        gw = std::max<int64_t>(lp.wpb, std::min<int64_t>(gw, (nq + lp.wpb - 1) / lp.wpb * lp.wpb));
        return (gw + lp.wpb - 1) / lp.wpb * lp.wpb;
    }
    int64_t slots = std::min<int64_t>(nq, std::max<int64_t>(lp.wpb, ((int64_t)1 << 30) >> lp.vlog2));
    slots = std::min<int64_t>(slots, 65536);
    return (slots + lp.wpb - 1) / lp.wpb * lp.wpb;
}

// Enqueue the search on `stream`.  vtab holds scratch_slots() tables of
// 2^vlog2 words (all zero; every query leaves its table zero), status nq
// words.  Enqueues kernels only, so it can be captured into a graph.
// Region order applies to persistent searches of whole batches that run the
// upper-layer descent on a graph with at least two upper levels.
static bool region_order(const gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, const KParams &P,
                         int64_t nq)
{
    return (p->flags & GVS_FLAG_REGION_ORDER) && p->persistent && !P.entries && !P.q_index && ix->max_level >= 2 &&
           lp.dfn && nq >= 2 * lp.wpb;
}

// entries, keys and permutation (double-buffered for the sort) for nq queries
static int ensure_region_scratch(gvs_index *ix, int64_t nq)
{
    if (nq > ix->rcap) {
        (void)hipFree(ix->d_rep);
        (void)hipFree(ix->d_rkey);
        (void)hipFree(ix->d_rperm);
        ix->d_rep = nullptr;
        ix->d_rkey = nullptr;
        ix->d_rperm = nullptr;
        ix->rcap = 0;
        if (!hip_ok(hipMalloc(&ix->d_rep, sizeof(int32_t) * nq), "hipMalloc region entries") ||
            !hip_ok(hipMalloc(&ix->d_rkey, 2 * sizeof(uint64_t) * nq), "hipMalloc region keys") ||
            !hip_ok(hipMalloc(&ix->d_rperm, 2 * sizeof(int32_t) * nq), "hipMalloc region order"))
            return GVS_ERR_NOMEM;
        ix->rcap = nq;
    }
    size_t bytes = 0;
    if (!hip_ok(rocprim::radix_sort_pairs(nullptr, bytes, ix->d_rkey, ix->d_rkey + ix->rcap, ix->d_rperm,
                                         ix->d_rperm + ix->rcap, (unsigned int)nq, 0, 62),
                "rocprim::radix_sort_pairs (size)"))
        return GVS_ERR_HIP;
    if (bytes > ix->rtmp_bytes) {
        (void)hipFree(ix->d_rtmp);
        ix->d_rtmp = nullptr;
        ix->rtmp_bytes = 0;
        if (!hip_ok(hipMalloc(&ix->d_rtmp, bytes), "hipMalloc sort storage")) return GVS_ERR_NOMEM;
        ix->rtmp_bytes = bytes;
    }
    return GVS_OK;
}

int launch_search(gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, KParams P, int64_t nq,
                  hipStream_t stream, uint32_t *vtab, int32_t *status, bool region)
{
    kfn_t fn = (kfn_t)lp.fn;
    // (with several waves per query, lds_per_wave is the whole workgroup's)
    const size_t shm = p->waves_per_query > 1 ? (size_t)lp.lds_per_wave : (size_t)lp.wpb * lp.lds_per_wave;
    const int64_t slots = scratch_slots(p, lp, nq);
    P.vtab = vtab;
    P.status = status;
    if (p->waves_per_query > 1 && !p->persistent) {
        // one workgroup of waves_per_query waves per query, chunks of `slots`
        P.persistent = 0;
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, ix->d_counters);
        for (int64_t b = 0; b < nq; b += slots) {
            const int64_t cnt = std::min(slots, nq - b);
            P.q_base = b;
            P.nq = b + cnt;
            hipLaunchKernelGGL(fn, dim3((unsigned)cnt), dim3(lp.wpb * WAVE), shm, stream, P);
            if (!hip_ok(hipGetLastError(), "launch k_search_mw")) return GVS_ERR_HIP;
        }
        return GVS_OK;
    }
    if (p->persistent) {
        P.persistent = 1;
        P.nq = nq;
        P.q_base = 0;
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, ix->d_counters);
        if (region) {
            // descent, then sort the batch by region, then layer 0 in that order;
            // the key is two node ids of ceil(log2 n) bits, so the radix sort
            // passes over only those bits
            P.key_bits = std::max(1, ilog2_ceil(std::max<int64_t>(ix->n, 2)));
            hipLaunchKernelGGL((dfn_t)lp.dfn, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, stream, P, ix->d_rep,
                               ix->d_rkey, ix->d_rperm);
            if (!hip_ok(hipGetLastError(), "launch k_descend")) return GVS_ERR_HIP;
            size_t bytes = ix->rtmp_bytes;
            if (!hip_ok(rocprim::radix_sort_pairs(ix->d_rtmp, bytes, ix->d_rkey, ix->d_rkey + ix->rcap, ix->d_rperm,
                                                 ix->d_rperm + ix->rcap, (unsigned int)nq, 0, 2 * P.key_bits,
                                                 stream),
                        "rocprim::radix_sort_pairs"))
                return GVS_ERR_HIP;
            P.entries = ix->d_rep;
            P.q_perm = ix->d_rperm + ix->rcap;
            P.keep_upper_stats = 1;
            P.xcd_queues = 1;
        }
        // a grid of resident waves, or of resident workgroups with several waves per query
        const int64_t blocks = p->waves_per_query > 1 ? slots : slots / lp.wpb;
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(lp.wpb * WAVE), shm, stream, P);
        if (!hip_ok(hipGetLastError(), "launch k_search (persistent)")) return GVS_ERR_HIP;
        return GVS_OK;
    }
    // one wave per query; queries processed in chunks of at most `slots`
    P.persistent = 0;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, ix->d_counters);
    for (int64_t b = 0; b < nq; b += slots) {
        int64_t cnt = std::min(slots, nq - b);
        P.q_base = b;
        P.nq = b + cnt;
        hipLaunchKernelGGL(fn, dim3((unsigned)((cnt + lp.wpb - 1) / lp.wpb)), dim3(lp.wpb * WAVE), shm, stream, P);
        if (!hip_ok(hipGetLastError(), "launch k_search")) return GVS_ERR_HIP;
    }
    return GVS_OK;
}

// launch_search on the index's own scratch, grown as needed
int launch_search_ix(gvs_index *ix, const gvs_params *p, const LaunchPlan &lp, const KParams &P, int64_t nq,
                     hipStream_t stream)
{
    int rc = ensure_scratch(ix, scratch_slots(p, lp, nq), lp.vlog2, nq);
    if (rc) return rc;
    const bool region = region_order(ix, p, lp, P, nq);
    if (region && (rc = ensure_region_scratch(ix, nq))) return rc;
    return launch_search(ix, p, lp, P, nq, stream, ix->d_vtab, ix->d_status, region);
}

// The most waves a batch of nq queries keeps busy: one per query, or the
// persistent grid when one is given.
static int64_t batch_waves(const gvs_params *p, int64_t nq)
{
    return p->persistent && p->grid_waves > 0 ? std::min<int64_t>(nq, p->grid_waves) : nq;
}

}  // namespace gvs

// An initial candidate pool per query (the NSG search's start): mode 1 =
// len ids per query row at ids + row * ld, mode 2 = the entry's first
// neighbours.
struct PoolInit {
    int mode = 0;
    const int32_t *ids = nullptr;
    int len = 0;
    int64_t ld = 0;
};

static int search_impl(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq, const int32_t *entries,
                       const gvs_params *p, int64_t *labels, float *dists, int32_t *stats, hipStream_t stream,
                       const PoolInit *pi = nullptr)
{
    if (!ix || !p || nq < 0 || (nq > 0 && (!queries || !labels || !dists)) || ldq < ix->dim) {
        gvs::set_error("gvs_search: invalid argument");
        return GVS_ERR_ARG;
    }
    if (pi && pi->mode) {
        if (entries || p->waves_per_query > 1 || (p->flags & GVS_FLAG_REGION_ORDER)) {
            gvs::set_error("gvs_search_pool: an initial pool excludes entries, several waves per query and "
                           "region order");
            return GVS_ERR_UNSUPPORTED;
        }
        if (pi->mode == 1 && (!pi->ids || pi->len < 1 || pi->len > std::max(p->ef, p->k) || pi->ld < pi->len)) {
            gvs::set_error("gvs_search_pool: need 1 <= init_len <= max(ef, k) and init_ld >= init_len");
            return GVS_ERR_ARG;
        }
    }
    if (nq == 0) return GVS_OK;
    if (nq > INT32_MAX) {
        gvs::set_error("gvs_search: nq too large");
        return GVS_ERR_ARG;
    }
    gvs::LaunchPlan lp;
    const bool pool = pi && pi->mode;
    int rc = gvs::plan_search(ix, p, p->ef, p->k, &lp, gvs::batch_waves(p, nq), pool);
    if (rc) return rc;
    KParams P;
    gvs::fill_params(ix, p, lp, P);
    P.queries = queries;
    P.ldq = ldq;
    P.entries = entries;
    P.k = p->k;
    P.ef = std::max(p->ef, p->k);
    P.out_labels = labels;
    P.out_dists = dists;
    P.out_stats = stats;
    if (pi && pi->mode) {
        P.init_mode = pi->mode;
        P.init_ids = pi->ids;
        P.init_len = pi->len;
        P.init_ld = pi->ld;
    }
    rc = gvs::launch_search_ix(ix, p, lp, P, nq, stream);
    if (rc) return rc;
    if (p->flags & GVS_FLAG_ASYNC) return GVS_OK;
    // visited-table overflow: rerun the affected queries with 4x larger tables
    uint32_t novf = 0;
    if (!hip_ok(hipMemcpyAsync(&novf, ix->d_counters + 1, 4, hipMemcpyDeviceToHost, stream), "read overflow count") ||
        !hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize"))
        return GVS_ERR_HIP;
    // idx[i] = the query that retry row i re-runs; a retry launch writes its
    // statuses by retry row, so each attempt keeps the rows that failed again
    std::vector<int64_t> idx;
    if (novf > 0) {
        std::vector<int32_t> st((size_t)nq);
        if (!hip_ok(hipMemcpy(st.data(), ix->d_status, sizeof(int32_t) * nq, hipMemcpyDeviceToHost), "read status"))
            return GVS_ERR_HIP;
        for (int64_t i = 0; i < nq; i++)
            if (st[(size_t)i]) idx.push_back(i);
    }
    for (int attempt = 0; !idx.empty() && attempt < 3; attempt++) {
        gvs_params p2 = *p;
        p2.visited_log2 = lp.vlog2 + 2 * (attempt + 1);
        p2.persistent = 0;
        p2.flags = (p->flags & ~GVS_FLAG_LDS_VISITED) | GVS_FLAG_HBM_VISITED;
        gvs::LaunchPlan lp2;
        rc = gvs::plan_search(ix, &p2, p2.ef, p2.k, &lp2, 0, pool);
        if (rc) return rc;
        KParams P2;
        gvs::fill_params(ix, &p2, lp2, P2);
        int64_t *d_idx = nullptr;
        if (!hip_ok(hipMalloc(&d_idx, sizeof(int64_t) * idx.size()), "hipMalloc retry") ||
            !hip_ok(hipMemcpy(d_idx, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice), "copy retry")) {
            (void)hipFree(d_idx);
            return GVS_ERR_HIP;
        }
        // rerun the failed rows, writing into temporaries, then scatter back
        const int64_t m = (int64_t)idx.size();
        int64_t *t_lab = nullptr;
        float *t_d = nullptr;
        int32_t *t_st = nullptr;
        int32_t *t_ent = nullptr;
        std::vector<int32_t> st2((size_t)m);
        bool ok = hip_ok(hipMalloc(&t_lab, sizeof(int64_t) * m * p->k), "hipMalloc") &&
                  hip_ok(hipMalloc(&t_d, sizeof(float) * m * p->k), "hipMalloc") &&
                  hip_ok(hipMalloc(&t_st, sizeof(int32_t) * m * GVS_STATS_ROW), "hipMalloc");
        if (ok && entries) {
            std::vector<int32_t> he((size_t)m);
            for (int64_t i = 0; i < m; i++)
                (void)hipMemcpy(&he[(size_t)i], entries + idx[(size_t)i], 4, hipMemcpyDeviceToHost);
            ok = hip_ok(hipMalloc(&t_ent, sizeof(int32_t) * m), "hipMalloc") &&
                 hip_ok(hipMemcpy(t_ent, he.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice), "copy");
        }
        if (ok) {
            P2.queries = queries;
            P2.ldq = ldq;
            P2.q_index = d_idx;
            P2.entries = t_ent;
            P2.k = p->k;
            P2.ef = std::max(p->ef, p->k);
            P2.out_labels = t_lab;
            P2.out_dists = t_d;
            P2.out_stats = t_st;
            // the pool rows follow the query rows (q_index)
            P2.init_mode = P.init_mode;
            P2.init_ids = P.init_ids;
            P2.init_len = P.init_len;
            P2.init_ld = P.init_ld;
            ok = gvs::launch_search_ix(ix, &p2, lp2, P2, m, stream) == GVS_OK &&
                 hip_ok(hipStreamSynchronize(stream), "sync retry") &&
                 hip_ok(hipMemcpy(st2.data(), ix->d_status, sizeof(int32_t) * m, hipMemcpyDeviceToHost),
                        "read retry status");
            for (int64_t i = 0; ok && i < m; i++) {
                if (st2[(size_t)i]) continue;  // failed again: its rows are written by a later attempt
                const int64_t q = idx[(size_t)i];
                ok = hip_ok(hipMemcpy(labels + q * p->k, t_lab + i * p->k, sizeof(int64_t) * p->k,
                                      hipMemcpyDeviceToDevice), "scatter") &&
                     hip_ok(hipMemcpy(dists + q * p->k, t_d + i * p->k, sizeof(float) * p->k,
                                      hipMemcpyDeviceToDevice), "scatter") &&
                     (!stats || hip_ok(hipMemcpy(stats + q * GVS_STATS_ROW, t_st + i * GVS_STATS_ROW,
                                                 sizeof(int32_t) * GVS_STATS_ROW, hipMemcpyDeviceToDevice),
                                       "scatter"));
            }
        }
        (void)hipFree(d_idx);
        (void)hipFree(t_lab);
        (void)hipFree(t_d);
        (void)hipFree(t_st);
        (void)hipFree(t_ent);
        if (!ok) return GVS_ERR_HIP;
        std::vector<int64_t> again;
        for (int64_t i = 0; i < m; i++)
            if (st2[(size_t)i]) again.push_back(idx[(size_t)i]);
        idx.swap(again);
    }
    novf = (uint32_t)idx.size();
    if (novf > 0) {
        gvs::set_error("gvs_search: %u queries overflowed the visited table", novf);
        return GVS_ERR_OVERFLOW;
    }
    return GVS_OK;
}

namespace gvs {
// The search for the GPU builder: device queries, entries (negative =
// descend), synchronous retries.
int search_rows(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq, const int32_t *entries,
                const gvs_params *p, int64_t *labels, float *dists, hipStream_t stream)
{
    return search_impl(ix, queries, nq, ldq, entries, p, labels, dists, nullptr, stream);
}
}  // namespace gvs

extern "C" int gvs_search_device(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                 const int32_t *entries, const gvs_params *p, int64_t *labels, float *dists,
                                 int32_t *stats, void *stream)
{
    return search_impl(ix, queries, nq, ldq, entries, p, labels, dists, stats, (hipStream_t)stream);
}

// gvs_search_device without any allocation: the persistent search with its
// visited tables and status words in the device's `scratch` array.  The grid
// is capped to the tables that fit; no retries (as GVS_FLAG_ASYNC).
extern "C" int gvs_search_device_noalloc(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                         const int32_t *entries, const gvs_params *p, int64_t *labels,
                                         float *dists, int32_t *stats, void *stream)
{
    if (!ix || !p || nq < 0 || (nq > 0 && (!queries || !labels || !dists)) || ldq < ix->dim) {
        gvs::set_error("gvs_search_device_noalloc: invalid argument");
        return GVS_ERR_ARG;
    }
    if (nq == 0) return GVS_OK;
    if (p->waves_per_query != 1) {
        gvs::set_error("gvs_search_device_noalloc: one wave per query only");
        return GVS_ERR_UNSUPPORTED;
    }
    int dev = -1;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return GVS_ERR_HIP;
    if (dev != ix->device) {
        gvs::set_error("gvs_search_device_noalloc: index is on device %d, current device is %d", ix->device, dev);
        return GVS_ERR_ARG;
    }
    unsigned char *base = nullptr;
    int rc = gvs::noalloc_scratch(dev, &base);
    if (rc) return rc;
    gvs_params q = *p;
    q.persistent = 1;
    q.flags = (p->flags | GVS_FLAG_ASYNC) & ~GVS_FLAG_REGION_ORDER;
    gvs::LaunchPlan lp;
    if ((rc = gvs::plan_search(ix, &q, q.ef, q.k, &lp, gvs::batch_waves(&q, nq)))) return rc;
    const size_t table_bytes = (size_t)4 << lp.vlog2;
    const size_t status_bytes = ((size_t)4 * (size_t)nq + 255) / 256 * 256;
    const size_t room = status_bytes < NOALLOC_SCRATCH_BYTES ? NOALLOC_SCRATCH_BYTES - status_bytes : 0;
    const int64_t fit = (int64_t)(room / table_bytes) / lp.wpb * lp.wpb;
    if (fit < lp.wpb) {
        gvs::set_error("gvs_search_device_noalloc: %lld queries with %zu-byte visited tables need more than the "
                       "%zu-byte scratch (at most %zu status bytes and %d tables)", (long long)nq, table_bytes,
                       NOALLOC_SCRATCH_BYTES, NOALLOC_SCRATCH_BYTES - (size_t)lp.wpb * table_bytes, lp.wpb);
        return GVS_ERR_ARG;
    }
    const int64_t want = q.grid_waves > 0 ? q.grid_waves : lp.resident_waves;
    q.grid_waves = (int)std::min<int64_t>(want, fit);
    // re-plan for the capped grid (the LDS hash size depends on the waves in use)
    if ((rc = gvs::plan_search(ix, &q, q.ef, q.k, &lp, gvs::batch_waves(&q, nq)))) return rc;
    const int64_t slots = gvs::scratch_slots(&q, lp, nq);
    if ((size_t)slots * ((size_t)4 << lp.vlog2) > room) {
        gvs::set_error("gvs_search_device_noalloc: visited tables do not fit the scratch");
        return GVS_ERR_ARG;
    }
    int32_t *status = (int32_t *)base;
    uint32_t *vtab = (uint32_t *)(base + status_bytes);
    KParams P;
    gvs::fill_params(ix, &q, lp, P);
    P.queries = queries;
    P.ldq = ldq;
    P.entries = entries;
    P.k = q.k;
    P.ef = std::max(q.ef, q.k);
    P.out_labels = labels;
    P.out_dists = dists;
    P.out_stats = stats;
    // the tables start empty (a fill kernel, so the sequence can be captured)
    const int64_t n16 = (int64_t)(((size_t)slots << lp.vlog2) / 4);
    hipLaunchKernelGGL(k_zero16, dim3((unsigned)std::min<int64_t>((n16 + 255) / 256, 4096)), dim3(256), 0,
                       (hipStream_t)stream, (uint4 *)vtab, n16);
    if (!hip_ok(hipGetLastError(), "launch k_zero16")) return GVS_ERR_HIP;
    return gvs::launch_search(ix, &q, lp, P, nq, (hipStream_t)stream, vtab, status, false);
}

extern "C" int gvs_stats_per_query(void) { return GVS_STATS_ROW; }

extern "C" int gvs_search_plan(gvs_index *ix, int64_t nq, const gvs_params *p, int32_t plan[8])
{
    if (!ix || !p || !plan || nq < 1) {
        gvs::set_error("gvs_search_plan: invalid argument");
        return GVS_ERR_ARG;
    }
    gvs::LaunchPlan lp;
    const int rc = gvs::plan_search(ix, p, p->ef, p->k, &lp, gvs::batch_waves(p, nq));
    if (rc) return rc;
    plan[0] = lp.hlog2;
    plan[1] = lp.hnarrow;
    plan[2] = lp.vlog2;
    plan[3] = lp.blog2;
    plan[4] = lp.lds_per_wave;
    plan[5] = lp.wpb;
    plan[6] = (int32_t)std::min<int64_t>(lp.resident_waves, INT32_MAX);
    plan[7] = lp.lcap;
    return GVS_OK;
}

extern "C" int gvs_hash_murmur2_keylen4_device(const uint32_t *keys, const uint32_t *seeds, uint32_t *out,
                                               int64_t n, void *stream)
{
    if (n < 0 || (n > 0 && (!keys || !seeds || !out)) || n > ((int64_t)1 << 40)) {
        gvs::set_error("gvs_hash_murmur2_keylen4_device: invalid argument");
        return GVS_ERR_ARG;
    }
    if (n == 0) return GVS_OK;
    hipLaunchKernelGGL(k_murmur2_k4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, keys,
                       seeds, out, n);
    return hip_ok(hipGetLastError(), "launch k_murmur2_k4") ? GVS_OK : GVS_ERR_HIP;
}

extern "C" int gvs_device_copy(void *dst, const void *src, int64_t bytes, void *stream)
{
    if (!dst || !src || bytes < 0 || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15)) {
        gvs::set_error("gvs_device_copy: 16-byte aligned pointers and sizes required");
        return GVS_ERR_ARG;
    }
    if (bytes == 0) return GVS_OK;
    int64_t n16 = bytes / 16;
    unsigned blocks = (unsigned)std::min<int64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst, n16);
    return hip_ok(hipGetLastError(), "launch k_copy16") ? GVS_OK : GVS_ERR_HIP;
}

extern "C" int gvs_search_overflow_count(gvs_index *ix, int64_t *count)
{
    if (!ix || !count) {
        gvs::set_error("gvs_search_overflow_count: invalid argument");
        return GVS_ERR_ARG;
    }
    uint32_t c = 0;
    if (!hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") ||
        !hip_ok(hipMemcpy(&c, ix->d_counters + 1, 4, hipMemcpyDeviceToHost), "read overflow count"))
        return GVS_ERR_HIP;
    *count = c;
    return GVS_OK;
}

extern "C" int gvs_search_counters(gvs_index *ix, int64_t counters[3])
{
    if (!ix || !counters) {
        gvs::set_error("gvs_search_counters: invalid argument");
        return GVS_ERR_ARG;
    }
    uint32_t c[4] = {0u, 0u, 0u, 0u};
    if (!hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") ||
        !hip_ok(hipMemcpy(c, ix->d_counters, sizeof c, hipMemcpyDeviceToHost), "read counters"))
        return GVS_ERR_HIP;
    for (int i = 0; i < 3; i++) counters[i] = c[i];
    return GVS_OK;
}

// ---- autotuner ----------------------------------------------------------
// Times the search of the caller's device queries over a grid of launch
// configurations and returns the fastest (the reference tunes (mc, mg) by
// exhaustive kernel timing, perf_test_scripts/perf_test_throughput.py:83-133,
// selection networked_FPGA/CPU_programs/launch_CPU_and_FPGA.py:172-183).
extern "C" int gvs_autotune(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq, const gvs_params *p,
                            int64_t *labels, float *dists, void *stream, gvs_params *best, float *best_ms)
{
    if (!ix || !p || !best || !best_ms || nq <= 0 || !queries || !labels || !dists || ldq < ix->dim) {
        gvs::set_error("gvs_autotune: invalid argument");
        return GVS_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const int rpc = (ix->r0 + WAVE - 1) / WAVE;
    std::vector<gvs_params> cand;
    for (int mc : {1, 2, 4, 8}) {
        if (mc * rpc > MAX_ROWS) continue;
        for (int mg : {1, 2, 4})
            for (int pers : {0, 1}) {
                gvs_params c = *p;
                c.mc = mc;
                c.mg = mg;
                c.persistent = pers;
                c.waves_per_query = 1;
                c.grid_waves = 0;
                c.flags = (p->flags | GVS_FLAG_ASYNC);
                cand.push_back(c);
                if (pers && ix->max_level >= 2 && nq >= 4096) {
                    gvs_params r = c;
                    r.flags |= GVS_FLAG_REGION_ORDER;
                    cand.push_back(r);
                }
                if (pers) {
                    gvs::LaunchPlan lp;
                    if (gvs::plan_search(ix, &c, c.ef, c.k, &lp) == GVS_OK && lp.resident_waves >= 2 * lp.wpb) {
                        c.grid_waves = lp.resident_waves / 2;  // half the resident waves
                        cand.push_back(c);
                    }
                }
            }
        // several waves per query: for few queries, every shape; for full
        // batches, two or four waves on the persistent queue (one query's
        // waves share its pool and visited hash, so a query needs less LDS
        // per wave and finishes sooner, which shortens the batch's tail)
        if (nq <= 64 && p->storage == GVS_STORE_F32)
            for (int W : {2, 4, 8})
                for (int mg : {1, 2, 4}) {
                    gvs_params c = *p;
                    c.mc = mc;
                    c.mg = mg;
                    c.persistent = 0;
                    c.waves_per_query = W;
                    c.grid_waves = 0;
                    c.flags = (p->flags | GVS_FLAG_ASYNC);
                    cand.push_back(c);
                }
        if (nq > 64 && mc >= 2 && p->storage == GVS_STORE_F32)
            for (int W : {2, 4}) {
                gvs_params c = *p;
                c.mc = mc;
                c.mg = 1;
                c.persistent = 1;
                c.waves_per_query = W;
                c.grid_waves = 0;
                c.flags = (p->flags | GVS_FLAG_ASYNC);
                cand.push_back(c);
                if (ix->max_level >= 2 && nq >= 4096) {
                    c.flags |= GVS_FLAG_REGION_ORDER;
                    cand.push_back(c);
                }
            }
    }
    // Above ef 128 the visited set's placement is timed too, unless the
    // caller fixed it: the LDS hash costs fewer instructions and HBM atomics
    // when many waves search at once, while the Bloom filter + HBM table
    // overlaps each id's check with its gather, which a small batch's few
    // waves feel more (on 1M x 128 at ef 192-512 the plain mc = 1 search of
    // one query runs about as fast with it as with 4-byte LDS slots, and of
    // up to 1,024 queries 4-18% faster than with 2-byte slots)
    if (std::max(p->ef, p->k) > 128 && !(p->flags & (GVS_FLAG_HBM_VISITED | GVS_FLAG_LDS_VISITED))) {
        const size_t n0 = cand.size();
        for (size_t i = 0; i < n0; i++)
            if (cand[i].waves_per_query == 1) {
                gvs_params h = cand[i];
                h.flags |= GVS_FLAG_HBM_VISITED;
                cand.push_back(h);
            }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")) {
        if (e0) (void)hipEventDestroy(e0);
        return GVS_ERR_HIP;
    }
    float best_t = -1.0f;
    int rc = GVS_OK;
    for (const gvs_params &c : cand) {
        gvs::LaunchPlan lp;
        if (gvs::plan_search(ix, &c, c.ef, c.k, &lp) != GVS_OK) continue;  // not compiled / unsupported
        // one warm-up launch, then three timed ones
        if ((rc = search_impl(ix, queries, nq, ldq, nullptr, &c, labels, dists, nullptr, st)) != GVS_OK) break;
        if (!hip_ok(hipEventRecord(e0, st), "hipEventRecord")) {
            rc = GVS_ERR_HIP;
            break;
        }
        for (int r = 0; r < 3 && rc == GVS_OK; r++) rc = search_impl(ix, queries, nq, ldq, nullptr, &c, labels, dists, nullptr, st);
        float ms = 0.0f;
        if (rc != GVS_OK || !hip_ok(hipEventRecord(e1, st), "hipEventRecord") ||
            !hip_ok(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) {
            rc = rc ? rc : GVS_ERR_HIP;
            break;
        }
        ms /= 3.0f;
        if (best_t < 0.0f || ms < best_t) {
            best_t = ms;
            *best = c;
            best->flags = (p->flags & ~GVS_FLAG_REGION_ORDER) | (c.flags & (GVS_FLAG_REGION_ORDER | GVS_FLAG_HBM_VISITED));
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_t < 0.0f) {
        gvs::set_error("gvs_autotune: no configuration is supported for this index");
        return GVS_ERR_UNSUPPORTED;
    }
    *best_ms = best_t;
    return GVS_OK;
}

// ---- captured launches --------------------------------------------------
struct gvs_search_graph {
    gvs_index *ix = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *vtab = nullptr;   // the graph's own visited tables (pointers are baked into the graph)
    int32_t *status = nullptr;
};

static void free_search_graph(gvs_search_graph *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->vtab);
    (void)hipFree(g->status);
    delete g;
}

extern "C" int gvs_search_graph_create(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                       const int32_t *entries, const gvs_params *p, int64_t *labels,
                                       float *dists, int32_t *stats, gvs_search_graph **out)
{
    if (!ix || !p || !out || nq <= 0 || nq > INT32_MAX || !queries || !labels || !dists || ldq < ix->dim) {
        gvs::set_error("gvs_search_graph_create: invalid argument");
        return GVS_ERR_ARG;
    }
    gvs::LaunchPlan lp;
    int rc = gvs::plan_search(ix, p, p->ef, p->k, &lp, gvs::batch_waves(p, nq));
    if (rc) return rc;
    KParams P;
    gvs::fill_params(ix, p, lp, P);
    P.queries = queries;
    P.ldq = ldq;
    P.entries = entries;
    P.k = p->k;
    P.ef = std::max(p->ef, p->k);
    P.out_labels = labels;
    P.out_dists = dists;
    P.out_stats = stats;
    gvs_search_graph *g = new gvs_search_graph();
    g->ix = ix;
    const size_t words = (size_t)gvs::scratch_slots(p, lp, nq) << lp.vlog2;
    hipStream_t cs = nullptr;
    bool ok = hip_ok(hipMalloc(&g->vtab, 4 * words), "hipMalloc visited tables") &&
              hip_ok(hipMemset(g->vtab, 0, 4 * words), "hipMemset visited tables") &&
              hip_ok(hipMalloc(&g->status, sizeof(int32_t) * nq), "hipMalloc status") &&
              hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags") &&
              hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (ok) {
        rc = gvs::launch_search(ix, p, lp, P, nq, cs, g->vtab, g->status, false);
        hipError_t e = hipStreamEndCapture(cs, &g->graph);
        ok = rc == GVS_OK && hip_ok(e, "hipStreamEndCapture") &&
             hip_ok(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    if (cs) (void)hipStreamDestroy(cs);
    if (!ok) {
        free_search_graph(g);
        return rc ? rc : GVS_ERR_HIP;
    }
    *out = g;
    return GVS_OK;
}

extern "C" int gvs_search_graph_launch(gvs_search_graph *g, void *stream)
{
    if (!g || !g->exec) {
        gvs::set_error("gvs_search_graph_launch: invalid argument");
        return GVS_ERR_ARG;
    }
    return hip_ok(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch") ? GVS_OK : GVS_ERR_HIP;
}

extern "C" int gvs_search_graph_destroy(gvs_search_graph *g)
{
    if (!g) return GVS_ERR_ARG;
    free_search_graph(g);
    return GVS_OK;
}

extern "C" int gvs_search_pool_device(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                      const int32_t *init_ids, int32_t init_len, int64_t init_ld,
                                      const gvs_params *p, int64_t *labels, float *dists, int32_t *stats,
                                      void *stream)
{
    PoolInit pi;
    pi.mode = init_ids ? 1 : 2;
    pi.ids = init_ids;
    pi.len = init_len;
    pi.ld = init_ld;
    return search_impl(ix, queries, nq, ldq, nullptr, p, labels, dists, stats, (hipStream_t)stream, &pi);
}

extern "C" int gvs_search_pool_host(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                                    const int32_t *init_ids, int32_t init_len, int64_t init_ld, const gvs_params *p,
                                    int64_t *labels, float *dists, int32_t *stats)
{
    if (!ix || !p || nq < 0 || (nq > 0 && (!queries || !labels || !dists)) || ldq < ix->dim || p->k < 1 ||
        (init_ids && (init_len < 1 || init_ld < init_len))) {
        gvs::set_error("gvs_search_pool_host: invalid argument");
        return GVS_ERR_ARG;
    }
    if (nq == 0) return GVS_OK;
    const int k = p->k;
    float *dq = nullptr;
    int32_t *di = nullptr, *ds = nullptr;
    int64_t *dl = nullptr;
    float *dd = nullptr;
    int rc = GVS_ERR_HIP;
    bool ok = hip_ok(hipMalloc(&dq, sizeof(float) * nq * ldq), "hipMalloc queries") &&
              hip_ok(hipMemcpy(dq, queries, sizeof(float) * nq * ldq, hipMemcpyHostToDevice), "copy queries") &&
              hip_ok(hipMalloc(&dl, sizeof(int64_t) * nq * k), "hipMalloc labels") &&
              hip_ok(hipMalloc(&dd, sizeof(float) * nq * k), "hipMalloc dists");
    if (ok && init_ids)
        ok = hip_ok(hipMalloc(&di, sizeof(int32_t) * nq * init_ld), "hipMalloc pool") &&
             hip_ok(hipMemcpy(di, init_ids, sizeof(int32_t) * nq * init_ld, hipMemcpyHostToDevice), "copy pool");
    if (ok && stats) ok = hip_ok(hipMalloc(&ds, sizeof(int32_t) * nq * GVS_STATS_ROW), "hipMalloc stats");
    if (ok) {
        rc = gvs_search_pool_device(ix, dq, nq, ldq, di, init_len, init_ld, p, dl, dd, ds, nullptr);
        if (rc == GVS_OK) {
            ok = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") &&
                 hip_ok(hipMemcpy(labels, dl, sizeof(int64_t) * nq * k, hipMemcpyDeviceToHost), "copy labels") &&
                 hip_ok(hipMemcpy(dists, dd, sizeof(float) * nq * k, hipMemcpyDeviceToHost), "copy dists") &&
                 (!stats || hip_ok(hipMemcpy(stats, ds, sizeof(int32_t) * nq * GVS_STATS_ROW, hipMemcpyDeviceToHost),
                                    "copy stats"));
            if (!ok) rc = GVS_ERR_HIP;
        }
    }
    (void)hipFree(dq);
    (void)hipFree(di);
    (void)hipFree(ds);
    (void)hipFree(dl);
    (void)hipFree(dd);
    return rc;
}

extern "C" int gvs_search_host(gvs_index *ix, const float *queries, int64_t nq, int64_t ldq,
                               const int32_t *entries, const gvs_params *p, int64_t *labels, float *dists,
                               int32_t *stats)
{
    if (!ix || !p || nq < 0 || (nq > 0 && (!queries || !labels || !dists)) || ldq < (ix ? ix->dim : 1)) {
        gvs::set_error("gvs_search_host: invalid argument");
        return GVS_ERR_ARG;
    }
    if (nq == 0) return GVS_OK;
    float *dq = nullptr;
    int32_t *de = nullptr, *ds = nullptr;
    int64_t *dl = nullptr;
    float *dd = nullptr;
    int rc = GVS_ERR_HIP;
    const int k = p->k;
    if (k < 1) {
        gvs::set_error("gvs_search_host: k must be >= 1");
        return GVS_ERR_ARG;
    }
    bool ok = hip_ok(hipMalloc(&dq, sizeof(float) * nq * ldq), "hipMalloc queries") &&
              hip_ok(hipMemcpy(dq, queries, sizeof(float) * nq * ldq, hipMemcpyHostToDevice), "copy queries") &&
              hip_ok(hipMalloc(&dl, sizeof(int64_t) * nq * k), "hipMalloc labels") &&
              hip_ok(hipMalloc(&dd, sizeof(float) * nq * k), "hipMalloc dists");
    if (ok && entries)
        ok = hip_ok(hipMalloc(&de, sizeof(int32_t) * nq), "hipMalloc entries") &&
             hip_ok(hipMemcpy(de, entries, sizeof(int32_t) * nq, hipMemcpyHostToDevice), "copy entries");
    if (ok && stats) ok = hip_ok(hipMalloc(&ds, sizeof(int32_t) * nq * GVS_STATS_ROW), "hipMalloc stats");
    if (ok) {
        rc = search_impl(ix, dq, nq, ldq, de, p, dl, dd, ds, nullptr);
        if (rc == GVS_OK) {
            ok = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize") &&
                 hip_ok(hipMemcpy(labels, dl, sizeof(int64_t) * nq * k, hipMemcpyDeviceToHost), "copy labels") &&
                 hip_ok(hipMemcpy(dists, dd, sizeof(float) * nq * k, hipMemcpyDeviceToHost), "copy dists") &&
                 (!stats || hip_ok(hipMemcpy(stats, ds, sizeof(int32_t) * nq * GVS_STATS_ROW, hipMemcpyDeviceToHost),
                                    "copy stats"));
            if (!ok) rc = GVS_ERR_HIP;
        }
    }
    (void)hipFree(dq);
    (void)hipFree(de);
    (void)hipFree(ds);
    (void)hipFree(dl);
    (void)hipFree(dd);
    return rc;
}
