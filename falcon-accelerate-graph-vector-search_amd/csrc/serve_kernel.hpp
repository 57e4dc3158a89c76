// Persistent serving kernel: waves claim published requests from a ring in
// pinned host memory, search them with each request's own k and ef, write
// the results back to pinned host memory, raise a per-slot done flag and
// append (position, overflow bit, slot) to a completion ring.
#pragma once
#include "search_kernels.hpp"

namespace gvsk {

// Requests are published on SERVE_QUEUES rings, each with its own tail
// (host) and head (device): a claim is a compare-and-swap on one ring's
// head, and waves racing for new work spread over the rings instead of
// serialising on one counter.
constexpr int SERVE_QUEUES = 16;
// heads in device memory, one per 128-byte line (16 uint64)
constexpr int SERVE_HEAD_STRIDE = 16;
// Device-memory mirror of the host's tails (uint64 words): [0, QUEUES) the
// tails (with SERVE_STOP), [16] the wall-clock tick of the last refresh.
constexpr int SERVE_MIRROR_TIME = 16;
constexpr int SERVE_MIRROR_WORDS = 32;
// An idle wave refreshes the mirror when it is older than this (100 MHz
// ticks): the host's tails are then read at most every 2 us in all.
constexpr uint64_t SERVE_REFRESH_TICKS = 200;

// Top bit of every ring's tail: leave the loop (one read gives both).
constexpr uint64_t SERVE_STOP = 1ull << 63;

// Control block in pinned (fine-grained) host memory.
struct ServeCtl {
    uint64_t tail[SERVE_QUEUES];  // host: requests published on each ring | SERVE_STOP (one 128-byte line)
    uint32_t err;                 // device: 1 + ring position of a request whose slot index was out of range
    uint32_t pad32;
    uint64_t pad[15];
};

struct ServeArgs {
    KParams P;               // out_labels/out_dists/out_stats: pinned, per slot, stride P.k
    ServeCtl *ctl;           // pinned
    const uint32_t *ring;    // pinned: slot of ring q's request i at ring[q * rcap + (i & (rcap-1))]
    uint32_t rcap;           // entries per ring, a power of two
    int dim;
    const float *qbuf;       // pinned: slot-major queries, stride ldq
    const int32_t *qk;       // pinned, per slot
    const int32_t *qef;      // pinned, per slot
    const uint64_t *qseq;    // pinned, per slot: sequence number of the request in the slot
    uint64_t *done;          // pinned, per slot: sequence number of the last completed request
    uint64_t *cring;         // pinned completion rings, one per request ring: entry i of ring q at
                             // cring[q * ccap + (i & (ccap-1))] = (i + 1) << 32 | overflow << 31 | slot
    uint32_t ccap;           // entries per completion ring, a power of two, >= rcap
    uint32_t nslots;         // slots (a ring entry names one below this)
    uint64_t *head;          // device: next request index to claim on ring q at head[q * SERVE_HEAD_STRIDE]
    uint64_t *mirror;        // device: SERVE_MIRROR_WORDS words, the tails as last read from the host
    uint64_t *served;        // device: requests of ring q completed (its completion ring position) at
                             // served[q * SERVE_HEAD_STRIDE]: one counter per ring, since same-address
                             // device-scope atomics from all XCDs serialise
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

// Per-wave (or per-workgroup) claim state.
struct ClaimState {
    int home;         // ring this claimer looks at first
    int steal;        // the other ring it looks at next
    bool keep;        // one of the first SERVE_KEEP waves (idle rule below)
    uint64_t seen_h;  // home ring's head at the last look
    uint64_t last;    // wall-clock tick of the last work
};

// One wave claims the next published request of its home ring, else of one
// other ring (a different one each time), or decides to leave: on the stop
// bit or after idle_ticks without work (every wave therefore ends even if
// the host goes away).  Published tails are read from the device-memory
// mirror; an idle wave that finds the mirror older than SERVE_REFRESH_TICKS
// wins the right to refresh it (a compare-and-swap on its time stamp) and
// reads all the host's tails in one 128-byte read across the bus, so the
// bus sees a fixed polling rate however many waves are idle, and no wave
// waits on another: any idle wave refreshes a stale mirror.  Idle waves
// back off exponentially (up to 32 sleeps of ~4 us) between looks.  The
// reads are relaxed: an acquire invalidates the XCD's L2 under the searches
// that run there (none is taken).  Returns (wave-uniform) 0 with the claimed
// ring q and position item, 1 on stop, 2 when idle too long.
__device__ uint32_t serve_claim(const ServeArgs &A, ClaimState &C, int &q, uint64_t &item)
{
    const int lane = lane_id();
    uint32_t state = 0u;
    int backoff = 0;
    q = C.home;
    item = 0ull;
    for (;;) {
        // refresh the mirror if it is stale (lanes 0..15 read the tails)
        const uint64_t now = wall_clock64();
        uint32_t refresh = 0u;
        if (lane == 0) {
            const uint64_t mt = __hip_atomic_load(A.mirror + SERVE_MIRROR_TIME, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
            if (now - mt > SERVE_REFRESH_TICKS &&
                atomicCAS((unsigned long long *)(A.mirror + SERVE_MIRROR_TIME), (unsigned long long)mt,
                          (unsigned long long)now) == mt)
                refresh = 1u;
        }
        if (__builtin_amdgcn_readfirstlane(refresh)) {
            if (lane < SERVE_QUEUES) {
                // tails (and the stop bit) only grow: a late refresher cannot
                // move the mirror back
                const uint64_t tw = ld_sys64r(&A.ctl->tail[lane]);
                atomicMax((unsigned long long *)(A.mirror + lane), (unsigned long long)tw);
            }
        }
        if (lane == 0) {
            bool got = false;
            for (int look = 0; look < 2 && !got && state == 0u; look++) {
                const int r = look == 0 ? C.home : (C.home + C.steal) % SERVE_QUEUES;
                const uint64_t tw = __hip_atomic_load(A.mirror + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (tw & SERVE_STOP) {
                    state = 1u;
                    break;
                }
                uint64_t h = __hip_atomic_load(A.head + (size_t)r * SERVE_HEAD_STRIDE, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                // Idle rule.  Claimers beyond the first SERVE_KEEP waves leave
                // after idle_ticks / 10 without work of their own; the first
                // SERVE_KEEP run their clock from the last claim by ANY wave on
                // their home ring, so they stay while requests flow and leave
                // together once the loop is idle.  The pool never shrinks below
                // SERVE_KEEP waves.
                if (look == 0 && C.keep && h != C.seen_h) {
                    C.seen_h = h;
                    C.last = now;
                }
                // claim: a failed compare-and-swap returns the head another
                // wave moved it to, tried at once while below the tail
                while (h < tw) {
                    const uint64_t o = atomicCAS((unsigned long long *)(A.head + (size_t)r * SERVE_HEAD_STRIDE),
                                                 (unsigned long long)h, (unsigned long long)(h + 1));
                    if (o == h) {
                        got = true;
                        item = h;
                        q = r;
                        break;
                    }
                    h = o;
                }
            }
            C.steal = C.steal % (SERVE_QUEUES - 1) + 1;
            if (!got && state == 0u && (int64_t)(now - C.last) > (C.keep ? A.idle_ticks : A.idle_ticks / 10))
                state = 2u;
            if (got) state = 3u;
        }
        const uint32_t st = __builtin_amdgcn_readfirstlane(state);
        if (st != 0u) {
            state = st == 3u ? 0u : st;
            break;
        }
        for (int i = 0; i < (1 << backoff); i++) __builtin_amdgcn_s_sleep(127);
        if (backoff < 5) backoff++;
    }
    C.steal = __builtin_amdgcn_readfirstlane(C.steal);
    q = __builtin_amdgcn_readfirstlane(q);
    item = readlane_u64(item, 0);
    return __builtin_amdgcn_readfirstlane(state);
}

// The claimed request's slot (wave-uniform), or -1 after reporting a corrupt
// ring entry.  Every field of a request is read with a system-scope load
// (fresh from host memory, not a cached copy), and each read depends on the
// claim, which depends on the mirror value that a refresher read from the
// host's published tail: no acquire fence is needed, and none is taken (a
// system-scope acquire invalidates this XCD's L2 under the searches running
// there).
__device__ int serve_slot(const ServeArgs &A, int q, uint64_t item)
{
    const size_t rpos = (size_t)q * A.rcap + (size_t)(item & (uint64_t)(A.rcap - 1u));
    const uint32_t slot = __builtin_amdgcn_readfirstlane(ld_sys32r(&A.ring[rpos]));
    if (slot >= A.nslots) {
        if (lane_id() == 0)
            __hip_atomic_store(&A.ctl->err, (uint32_t)rpos + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return -1;
    }
    return (int)slot;
}

// The results were written by this wave with system-scope stores, which
// write through to host memory: once they are complete (the wait below,
// also a compiler barrier) the flags can follow, with no release fence (at
// system scope that writes back the whole L2 of the XCD, once per request).
// Then the per-slot done word and the entry in the completion ring of the
// request's ring q (position, overflow bit, slot), also stored through to
// host memory; the host reads the entry before the rows, and PCIe keeps
// posted writes in order.
__device__ void serve_complete(const ServeArgs &A, int q, int slot, uint64_t seq, bool ovf)
{
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (lane_id() == 0) {
        __hip_atomic_store(&A.done[slot], seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        const uint64_t pos = atomicAdd((unsigned long long *)(A.served + (size_t)q * SERVE_HEAD_STRIDE), 1ull);
        __hip_atomic_store(&A.cring[(size_t)q * A.ccap + (size_t)(pos & (uint64_t)(A.ccap - 1u))],
                           ((pos + 1ull) << 32) | ((uint64_t)(ovf ? 1u : 0u) << 31) | (uint64_t)slot,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// One wave per request.
template <int NCH, int STORE, int METRIC, int MROWS>
__global__ __launch_bounds__(256, GVS_MIN_WAVES_PER_EU) void k_serve(ServeArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wib = threadIdx.x >> 6;
    unsigned char *lds = smem + (size_t)wib * A.P.lds_per_wave;
    const int64_t gw = (int64_t)blockIdx.x * A.P.waves_per_block + wib;
    ClaimState C{(int)(gw % SERVE_QUEUES), 1, gw < SERVE_KEEP, ~0ull, (uint64_t)wall_clock64()};
    for (;;) {
        int q;
        uint64_t item;
        if (serve_claim(A, C, q, item) != 0u) break;
        const int slot = serve_slot(A, q, item);
        if (slot < 0) break;
        const int k = ld_sys32i(&A.qk[slot]);
        const int ef = ld_sys32i(&A.qef[slot]);
        const uint64_t seq = ld_sys64r(&A.qseq[slot]);
        const bool ovf = search_query_req<NCH, STORE, METRIC, MROWS>(A.P, (int64_t)slot, gw, lds, k, ef,
                                                                     A.qbuf + (size_t)slot * A.P.ldq, true);
        serve_complete(A, q, slot, seq, ovf);
        C.last = wall_clock64();
    }
}

// One workgroup of A.P.waves_per_block waves per request (the intra-query
// search, DST(mc, 1)): wave 0 claims, the workgroup searches, wave 0
// completes.  The workgroup's waves wait for wave 0 only at workgroup
// barriers (all waves of a workgroup are resident together).
template <int NCH, int STORE, int METRIC, int WROWS>
__global__ __launch_bounds__(512) void k_serve_mw(ServeArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int64_t claim[4];  // state, slot, k | ef << 32, seq
    const int w = threadIdx.x >> 6;
    const int W = A.P.waves_per_block;
    const int64_t gb = blockIdx.x;
    ClaimState C{(int)(gb % SERVE_QUEUES), 1, gb * W < SERVE_KEEP, ~0ull, (uint64_t)wall_clock64()};
    int q = 0;  // wave 0: the ring of the request in hand
    for (;;) {
        if (w == 0) {
            uint64_t item;
            uint32_t st = serve_claim(A, C, q, item);
            int slot = -1;
            int k = 0, ef = 0;
            uint64_t seq = 0ull;
            if (st == 0u) {
                slot = serve_slot(A, q, item);
                if (slot < 0) {
                    st = 4u;
                } else {
                    k = ld_sys32i(&A.qk[slot]);
                    ef = ld_sys32i(&A.qef[slot]);
                    seq = ld_sys64r(&A.qseq[slot]);
                }
            }
            if (lane_id() == 0) {
                claim[0] = st;
                claim[1] = slot;
                claim[2] = (int64_t)(uint32_t)k | ((int64_t)ef << 32);
                claim[3] = (int64_t)seq;
            }
        }
        wg_sync();
        const uint32_t st = (uint32_t)claim[0];
        const int slot = (int)claim[1];
        const int k = (int)(uint32_t)claim[2];
        const int ef = (int)(claim[2] >> 32);
        const uint64_t seq = (uint64_t)claim[3];
        if (st != 0u) break;  // the same for every wave of the workgroup
        // (the search's own barriers order these reads before the next claim's writes)
        const bool ovf = search_query_mw_req<NCH, STORE, METRIC, WROWS>(A.P, (int64_t)slot, gb, smem, k, ef,
                                                                        A.qbuf + (size_t)slot * A.P.ldq, true);
        if (w == 0) serve_complete(A, q, slot, seq, ovf);
        C.last = wall_clock64();
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

// rows per wave = ceil(mc * ceil(r0/64) / waves) rounded up to 1, 2 or 4
template <int NCH, int STORE, int METRIC>
sfn_t pick_serve_rows_mw(int rows_per_wave)
{
    if (rows_per_wave <= 1) return k_serve_mw<NCH, STORE, METRIC, 1>;
    if (rows_per_wave <= 2) return k_serve_mw<NCH, STORE, METRIC, 2>;
    if (rows_per_wave <= 4) return k_serve_mw<NCH, STORE, METRIC, 4>;
    return nullptr;
}

}  // namespace gvsk

gvsk::sfn_t gvs_serve_kernel_f32_l2(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_f32_ip(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_mw_f32_l2(int nch, int rows_per_wave);
gvsk::sfn_t gvs_serve_kernel_mw_f32_ip(int nch, int rows_per_wave);
gvsk::sfn_t gvs_serve_kernel_bf16_l2(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_bf16_ip(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_fp8_l2(int nch, int rows);
gvsk::sfn_t gvs_serve_kernel_fp8_ip(int nch, int rows);
