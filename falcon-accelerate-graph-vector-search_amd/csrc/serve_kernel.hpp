// Persistent serving kernel: waves claim published requests from a ring in
// pinned host memory, search them with each request's own k and ef, write
// the results back to pinned host memory and raise a per-slot done flag.
#pragma once
#include "search_kernels.hpp"

namespace gvsk {

// Control block in pinned (fine-grained) host memory.
struct ServeCtl {
    uint64_t tail;   // host: number of requests published
    uint32_t stop;   // host: 1 = leave the loop
    uint32_t pad;
};

struct ServeArgs {
    KParams P;               // out_labels/out_dists/out_stats: pinned, per slot, stride P.k
    ServeCtl *ctl;           // pinned
    const uint32_t *ring;    // pinned: slot of request i at ring[i & (rcap-1)]
    uint32_t rcap;           // power of two
    int dim;
    const float *qbuf;       // pinned: slot-major queries, stride ldq
    const int32_t *qk;       // pinned, per slot
    const int32_t *qef;      // pinned, per slot
    const uint64_t *qseq;    // pinned, per slot: sequence number of the request in the slot
    uint64_t *done;          // pinned, per slot: sequence number of the last completed request
    uint64_t *cring;         // pinned completion ring: entry i = (i + 1) << 32 | slot of the i-th completion
    uint32_t ccap;           // power of two, >= slots
    uint64_t *head;          // device: next request index to claim
    uint64_t *served;        // device: completed requests (completion ring position)
    int64_t idle_ticks;      // 100 MHz ticks without work before a wave leaves the loop
};

__device__ __forceinline__ uint64_t ld_sys64(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_sys32(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int32_t ld_sys32i(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// relaxed system-scope loads: read memory (not a cached copy) without the
// L2 invalidation an acquire brings
__device__ __forceinline__ uint64_t ld_sys64r(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_sys32r(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// waves that stay while requests flow (one per CU on MI355X)
constexpr int64_t SERVE_KEEP = 256;

#ifndef GVS_MIN_WAVES_PER_EU
#define GVS_MIN_WAVES_PER_EU 1
#endif

template <int NCH, int STORE, int METRIC, int MROWS>
__global__ __launch_bounds__(256, GVS_MIN_WAVES_PER_EU) void k_serve(ServeArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wib = threadIdx.x >> 6;
    unsigned char *lds = smem + (size_t)wib * A.P.lds_per_wave;
    const int64_t gw = (int64_t)blockIdx.x * A.P.waves_per_block + wib;
    const int lane = lane_id();
    uint64_t last = wall_clock64();
    for (;;) {
        // lane 0 claims the next published request, or decides to leave:
        // on the stop flag or after idle_ticks without work (every wave
        // therefore ends even if the host goes away)
        uint32_t state = 0u;
        uint64_t item = 0ull;
        if (lane == 0) {
            // Idle waves poll the host ring with exponential back-off (up to
            // 32 sleeps of ~4 us): every idle wave's poll is a read across
            // the bus, and thousands of them at full rate load the host's
            // memory; with many waves some wave still looks every few ns.
            // The polls are relaxed: an acquire invalidates the XCD's L2
            // under the searches that run there, so it is taken once per
            // claimed request (below), not once per poll.
            int backoff = 0;
            uint64_t seen_h = ~0ull;
            uint64_t h_max = 0ull;  // the highest head seen (it only grows)
            for (;;) {
                if (ld_sys32r(&A.ctl->stop)) {
                    state = 1u;
                    break;
                }
                const uint64_t t = ld_sys64r(&A.ctl->tail);
                // the local L2 may hold an old head: never go below one seen
                uint64_t h = __hip_atomic_load(A.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                h = h > h_max ? h : h_max;
                // Idle rule.  Waves gw >= SERVE_KEEP leave after
                // idle_ticks / 10 without work of their own; the first
                // SERVE_KEEP waves run their clock from the last claim by
                // ANY wave, so they stay while requests flow and leave
                // together once the loop is idle.  The pool never shrinks
                // below SERVE_KEEP waves.
                if (gw < SERVE_KEEP && h != seen_h) {
                    seen_h = h;
                    last = wall_clock64();
                }
                if (h < t) {
                    // claim: a failed compare-and-swap returns the head another
                    // wave moved it to, which is tried at once while it is
                    // below the tail read above (the tail only grows)
                    bool got = false;
                    while (h < t) {
                        const uint64_t o = atomicCAS((unsigned long long *)A.head, (unsigned long long)h,
                                                     (unsigned long long)(h + 1));
                        if (o == h) {
                            got = true;
                            break;
                        }
                        h = o;
                    }
                    if (got) {
                        item = h;
                        break;
                    }
                    h_max = h;  // every request up to the tail was claimed: back off
                }
                if ((int64_t)(wall_clock64() - last) > (gw < SERVE_KEEP ? A.idle_ticks : A.idle_ticks / 10)) {
                    state = 2u;
                    break;
                }
                for (int i = 0; i < (1 << backoff); i++) __builtin_amdgcn_s_sleep(127);
                if (backoff < 5) backoff++;
            }
        }
        state = __builtin_amdgcn_readfirstlane(state);
        if (state != 0u) break;
        item = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(item >> 32)) << 32) |
               (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)item);
        // The tail read (relaxed) showed the host's release of this request:
        // an acquire fence orders every read of its fields after it.
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        const uint32_t slot = ld_sys32r(&A.ring[item & (uint64_t)(A.rcap - 1u)]);
        const int k = ld_sys32i(&A.qk[slot]);
        const int ef = ld_sys32i(&A.qef[slot]);
        const uint64_t seq = ld_sys64r(&A.qseq[slot]);
        search_query_req<NCH, STORE, METRIC, MROWS>(A.P, (int64_t)slot, gw, lds, k, ef,
                                                    A.qbuf + (size_t)slot * A.P.ldq, true);
        // results (plain stores to pinned memory) become visible before the
        // flag: a release fence (write-back, no L2 invalidation; the next
        // claim takes its own acquire)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        if (lane == 0) {
            __hip_atomic_store(&A.done[slot], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t pos = atomicAdd((unsigned long long *)A.served, 1ull);
            __hip_atomic_store(&A.cring[pos & (uint64_t)(A.ccap - 1u)], ((pos + 1ull) << 32) | (uint64_t)slot,
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        last = wall_clock64();
    }
}

typedef void (*sfn_t)(ServeArgs);

template <int NCH, int STORE, int METRIC>
sfn_t pick_serve_rows(int rows)
{
    if (rows <= 1) return k_serve<NCH, STORE, METRIC, 1>;
    if (rows <= 2) return k_serve<NCH, STORE, METRIC, 2>;
    return k_serve<NCH, STORE, METRIC, 8>;
}

}  // namespace gvsk

gvsk::sfn_t gvs_serve_kernel_f32_l2(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_f32_ip(int nch, int rows);
