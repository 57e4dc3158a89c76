// One query on one wavefront: upper-layer greedy descent, then layer-0
// best-first search with DST groups of up to mc candidates per step.
#pragma once
#include "search_kernel.hpp"

namespace gvsk {

// LDS carve-up of one wave's partition (sizes in P).
struct WaveLds {
    uint64_t *poolA, *poolB, *lkeys;
    uint32_t *bloom, *lids, *sel;
};

__device__ __forceinline__ WaveLds carve(const KParams &P, unsigned char *base)
{
    WaveLds L;
    L.poolA = (uint64_t *)base;
    L.poolB = L.poolA + P.ef_cap;
    L.lkeys = L.poolB + P.ef_cap;
    L.bloom = (uint32_t *)(L.lkeys + P.lcap);
    L.lids = L.bloom + (1u << P.blog2);
    L.sel = L.lids + P.lcap;
    return L;
}

__device__ __forceinline__ uint32_t bloom_mask(uint32_t h1)
{
    return (1u << (h1 & 31u)) | (1u << ((h1 >> 5) & 31u)) | (1u << ((h1 >> 10) & 31u));
}

// Merge keys[0..n) (EMPTY entries allowed) into the sorted pool, keeping the
// ef smallest; returns the new pool size.  pool/pool2 are swapped.
__device__ __forceinline__ int merge_into_pool(uint64_t *&pool, uint64_t *&pool2, int np, int ef,
                                               uint64_t *keys, int n)
{
    const int lane = lane_id();
    const uint64_t thr = (np >= ef) ? pool[ef - 1] : EMPTY_KEY;
    int ns = 0;
    for (int b = 0; b < n; b += WAVE) {
        int i = b + lane;
        uint64_t key = i < n ? keys[i] : EMPTY_KEY;
        bool s = key < thr;
        uint64_t bal = __ballot(s);
        wave_sync();
        if (s) keys[ns + mbcnt(bal)] = key;
        ns += __popcll(bal);
    }
    wave_sync();
    if (ns == 0) return np;
    for (int i = lane; i < ns; i += WAVE) {
        uint64_t s = keys[i];
        int r = 0;
        for (int j = 0; j < ns; j++) r += (keys[j] < s) ? 1 : 0;
        int lo = 0, hi = np;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (pool[mid] < s) lo = mid + 1;
            else hi = mid;
        }
        int dest = r + lo;
        if (dest < ef) pool2[dest] = s;
    }
    for (int i = lane; i < np; i += WAVE) {
        uint64_t v = pool[i];
        int c = 0;
        for (int j = 0; j < ns; j++) c += (keys[j] < v) ? 1 : 0;
        int dest = i + c;
        if (dest < ef) pool2[dest] = v;
    }
    wave_sync();
    uint64_t *t = pool;
    pool = pool2;
    pool2 = t;
    return min(ef, np + ns);
}

template <int NCH, int STORE, int METRIC>
__device__ void search_query(const KParams &P, int64_t qi, int64_t slot, unsigned char *lds)
{
    constexpr int E = StoreTraits<STORE>::E;
    const int lane = lane_id();
    const int p = lane & 7;
    WaveLds L = carve(P, lds);
    const uint32_t nbw = 1u << P.blog2;

    const int64_t qrow = P.q_index ? P.q_index[qi] : qi;
    const int k = P.q_k ? P.q_k[qi] : P.k;
    int ef = P.q_ef ? P.q_ef[qi] : P.ef;
    if (ef < k) ef = k;

    float qf[NCH * E];
    const float *Q = P.queries + qrow * P.ldq;
#pragma unroll
    for (int i = 0; i < NCH; i++)
#pragma unroll
        for (int t = 0; t < E; t++) {
            int idx = (p + 8 * i) * E + t;
            qf[i * E + t] = idx < P.dim ? Q[idx] : 0.0f;
        }

    for (uint32_t w = lane; w < nbw; w += WAVE) L.bloom[w] = 0u;

    int up_hops = 0, up_evals = 0;
    uint32_t ep;
    float epd = 0.0f;
    bool have_epd = false;
    if (P.entries) {
        ep = (uint32_t)P.entries[qi];
    } else {
        // upper-layer greedy descent, HNSWLIB/hnswalg.h:1119-1147
        uint32_t cur = (uint32_t)P.entry;
        if (lane == 0) L.lids[0] = cur;
        wave_sync();
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, 1, L.lkeys);
        wave_sync();
        float curd = key_dist(L.lkeys[0]);
        up_evals = 1;
        for (int level = P.max_level; level > 0; level--) {
            for (;;) {
                const int64_t row = P.upper_off[cur] + (level - 1);
                int32_t nid = lane < P.ru ? P.adj_up[row * P.ru + lane] : -1;
                uint64_t bal = __ballot(nid >= 0);
                int n = __popcll(bal);
                wave_sync();
                if (nid >= 0) L.lids[mbcnt(bal)] = (uint32_t)nid;
                wave_sync();
                eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, n, L.lkeys);
                wave_sync();
                up_hops++;
                up_evals += n;
                uint64_t best = EMPTY_KEY;
                if (lane < n) best = ((uint64_t)ord_f(key_dist(L.lkeys[lane])) << 32) | (uint64_t)lane;
                best = wave_min_u64(best);
                if (n > 0) {
                    float bd = unord_f((uint32_t)(best >> 32));
                    if (bd < curd) {   // strictly closer
                        curd = bd;
                        cur = L.lids[(int)(best & 63u)];
                        wave_sync();
                        continue;
                    }
                }
                break;
            }
        }
        ep = cur;
        epd = curd;
        have_epd = true;
    }

    if (!have_epd) {
        if (lane == 0) L.lids[0] = ep;
        wave_sync();
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, 1, L.lkeys);
        wave_sync();
        epd = key_dist(L.lkeys[0]);
    }

    // layer-0 best-first search, HNSWLIB/hnswalg.h:240-326, DST groups as
    // INTER/scheduler.hpp:100-157 with one group in flight
    uint64_t *pool = L.poolA, *pool2 = L.poolB;
    int np = 1;
    uint32_t *T = P.vtab + ((size_t)slot << P.vlog2);
    const uint32_t tmask = (1u << P.vlog2) - 1u;
    const int tlimit = (int)((3u << P.vlog2) >> 2) - P.lcap;
    wave_sync();
    if (lane == 0) {
        pool[0] = mk_key(epd, ep);
        uint32_t h0 = murmur2_k4(ep, P.bloom_seed), h1 = murmur2_k4(ep, P.bloom_seed + 1u);
        L.bloom[h0 & (nbw - 1u)] |= bloom_mask(h1);
        uint32_t pos = (ep * 0x9E3779B1u) >> (32 - P.vlog2);
        atomicCAS(&T[pos], 0u, ep + 1u);
    }
    int hops = 0, evals = 1;
    bool overflow = false;
    const int rpc = (P.r0 + WAVE - 1) / WAVE;
    wave_sync();

    for (;;) {
        // pick up to mc unexpanded entries, closest first
        int nsel = 0;
        for (int base = 0; base < np && nsel < P.mc; base += WAVE) {
            int i = base + lane;
            uint64_t key = i < np ? pool[i] : EMPTY_KEY;
            bool unexp = (i < np) && !(key & 1ull);
            uint64_t bal = __ballot(unexp);
            while (bal && nsel < P.mc) {
                int l = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                if (lane == l) {
                    pool[i] = key | 1ull;
                    L.sel[nsel] = key_id(key);
                }
                nsel++;
            }
        }
        if (nsel == 0) break;
        wave_sync();
        hops += nsel;
        const int nrows = nsel * rpc;

        int32_t ids[MAX_ROWS];
        uint32_t prev[MAX_ROWS], pos[MAX_ROWS];
        int spos[MAX_ROWS];
#pragma unroll
        for (int r = 0; r < MAX_ROWS; r++) {
            ids[r] = -1;
            prev[r] = 0u;
            pos[r] = 0u;
            spos[r] = 0;
            if (r < nrows) {
                int ci = r / rpc, part = r - ci * rpc;
                int j = part * WAVE + lane;
                if (j < P.r0) ids[r] = P.adj0[(size_t)L.sel[ci] * P.r0 + j];
            }
        }
        // Bloom probe (LDS) + exact visited insert (HBM CAS) + speculative list
        uint32_t spec_bits = 0u;
        int nl = 0;
#pragma unroll
        for (int r = 0; r < MAX_ROWS; r++) {
            if (r < nrows) {
                const int32_t id = ids[r];
                const bool valid = id >= 0;
                bool maybe = true;
                if (valid) {
                    uint32_t h0 = murmur2_k4((uint32_t)id, P.bloom_seed);
                    uint32_t h1 = murmur2_k4((uint32_t)id, P.bloom_seed + 1u);
                    uint32_t m = bloom_mask(h1);
                    uint32_t old = atomicOr(&L.bloom[h0 & (nbw - 1u)], m);
                    maybe = (old & m) == m;
                    pos[r] = ((uint32_t)id * 0x9E3779B1u) >> (32 - P.vlog2);
                    prev[r] = atomicCAS(&T[pos[r]], 0u, (uint32_t)id + 1u);
                }
                const bool spec = valid && !maybe;
                uint64_t bal = __ballot(spec);
                if (spec) {
                    spos[r] = nl + mbcnt(bal);
                    L.lids[spos[r]] = (uint32_t)id;
                }
                nl += __popcll(bal);
                spec_bits |= (spec ? 1u : 0u) << r;
            }
        }
        wave_sync();
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, nl, L.lkeys);
        // resolve the CAS results
        int nl2 = nl;
#pragma unroll
        for (int r = 0; r < MAX_ROWS; r++) {
            if (r < nrows) {
                const int32_t id = ids[r];
                const bool valid = id >= 0;
                const uint32_t want = (uint32_t)id + 1u;
                bool pending = valid && prev[r] != 0u && prev[r] != want;
                while (__ballot(pending)) {
                    if (pending) {
                        pos[r] = (pos[r] + 1u) & tmask;
                        prev[r] = atomicCAS(&T[pos[r]], 0u, want);
                        pending = prev[r] != 0u && prev[r] != want;
                    }
                }
                const bool won = valid && prev[r] == 0u;
                const bool spec = (spec_bits >> r) & 1u;
                if (spec && !won) L.lkeys[spos[r]] = EMPTY_KEY;   // same id twice in one group
                const bool late = won && !spec;                    // Bloom false positive
                uint64_t bal = __ballot(late);
                if (late) L.lids[nl2 + mbcnt(bal)] = (uint32_t)id;
                nl2 += __popcll(bal);
                evals += __popcll(__ballot(won));
            }
        }
        wave_sync();
        if (nl2 > nl) {
            eval_rows<NCH, STORE, METRIC>(P, qf, L.lids + nl, nl2 - nl, L.lkeys + nl);
            wave_sync();
        }
        if (evals > tlimit) {
            overflow = true;
            break;
        }
        np = merge_into_pool(pool, pool2, np, ef, L.lkeys, nl2);
    }

    // results, closer first (HNSWLIB/hnswlib.h:171-185)
    for (int i = lane; i < P.k; i += WAVE) {
        uint64_t key = (!overflow && i < np && i < k) ? pool[i] : EMPTY_KEY;
        int64_t lab = -1;
        float d = __builtin_inff();
        if (key != EMPTY_KEY) {
            uint32_t id = key_id(key);
            lab = P.labels ? P.labels[id] : (int64_t)id;
            d = key_dist(key);
        }
        P.out_labels[qi * P.k + i] = lab;
        P.out_dists[qi * P.k + i] = d;
    }
    if (lane == 0) {
        if (P.out_stats) {
            P.out_stats[qi * 4 + 0] = hops;
            P.out_stats[qi * 4 + 1] = evals;
            P.out_stats[qi * 4 + 2] = up_hops;
            P.out_stats[qi * 4 + 3] = up_evals;
        }
        if (P.status) P.status[qi] = overflow ? 1 : 0;
        if (overflow) atomicAdd(&P.counters[1], 1u);
    }
    // leave the slot's visited table empty for the next query
    uint4 *T4 = (uint4 *)T;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t i = lane; i < ((1u << P.vlog2) >> 2); i += WAVE) T4[i] = z;
    __threadfence();
    wave_sync();
}

}  // namespace gvsk
