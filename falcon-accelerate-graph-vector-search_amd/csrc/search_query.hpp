// One query on one wavefront: upper-layer greedy descent, then layer-0
// best-first search with DST groups (mc candidates per group, mg groups in
// flight).
#pragma once
#include "search_kernel.hpp"

namespace gvsk {

// Optional phase timing (built only with -DGVS_PROFILE): per query, the
// stats row holds 16 int32: the 4 counters, then phase cycles / 64.
#ifdef GVS_PROFILE
#define PROF_DECL uint64_t prof_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}; uint64_t prof_t = clock64();
#define PROF_MARK(ph) do { uint64_t _n = clock64(); prof_acc[ph] += _n - prof_t; prof_t = _n; } while (0)
#else
#define PROF_DECL
#define PROF_MARK(ph) do { } while (0)
#endif

// LDS carve-up of one wave's partition (sizes in P).
struct WaveLds {
    uint64_t *poolA, *poolB, *lkeys, *bloom;
    uint32_t *lids, *sel, *vh;  // sel: ring of selected groups, 8 ids per group
    uint32_t nbw;               // Bloom words
};

__device__ __forceinline__ WaveLds carve(const KParams &P, unsigned char *base)
{
    WaveLds L;
    L.poolA = (uint64_t *)base;
    L.poolB = L.poolA + P.ef_cap;
    L.lkeys = L.poolB + P.ef_cap;
    L.bloom = L.lkeys + P.lcap;
    L.lids = (uint32_t *)(L.bloom + (1u << P.blog2));
    L.sel = L.lids + P.lcap;
    L.vh = L.sel + gvs_sel_words(P.mg);
    L.nbw = 1u << P.blog2;
    return L;
}

// Blocked Bloom filter: one 64-bit word per id, six bit positions.
// word = h0 mod nwords; bits from h1 (five 6-bit fields) and the top of h0.
__device__ __forceinline__ void bloom_pos(uint32_t id, uint32_t seed, uint32_t nwords, uint32_t &w, uint64_t &m)
{
    uint32_t h0 = murmur2_k4(id, seed), h1 = murmur2_k4(id, seed + 1u);
    w = h0 & (nwords - 1u);
    m = (1ull << (h1 & 63u)) | (1ull << ((h1 >> 6) & 63u)) | (1ull << ((h1 >> 12) & 63u)) |
        (1ull << ((h1 >> 18) & 63u)) | (1ull << ((h1 >> 24) & 63u)) | (1ull << (h0 >> 26));
}

__device__ __forceinline__ uint32_t vhash(uint32_t id, int vlog2) { return (id * 0x9E3779B1u) >> (32 - vlog2); }

// ---- 2-byte LDS hash slots (ef above 128) ------------------------------
// For graphs of at most 2^(hlog2 + 8) nodes.  y = id * odd mod 2^hbits is a
// bijection of the ids; its top hlog2 bits are the node's home slot, the
// other hbits - hlog2 (<= 8) bits its quotient.  A node stored at probe step
// d (the same triangular sequence as the 4-byte tables) holds
// 1 + (quotient << 7 | d).  A slot whose step and quotient both match is the
// node itself (same home, same quotient, same y), so membership stays exact.
// A probe sequence past step 127 (probability about load^128) makes the
// query overflow.  Slots pair up in 32-bit words, claimed with a 32-bit LDS
// compare-and-swap.
__device__ __forceinline__ void nkey(uint32_t id, int hbits, int hlog2, uint32_t &home, uint32_t &q)
{
    const uint32_t y = (id * 0x9E3779B1u) & ((1u << hbits) - 1u);
    home = y >> (hbits - hlog2);
    q = y & ((1u << (hbits - hlog2)) - 1u);
}
__device__ __forceinline__ uint32_t nentry(uint32_t q, uint32_t d) { return 1u + ((q << 7) | d); }

// Pick up to mc unexpanded pool entries, closest first, from index `hint`
// on (entries before hint are expanded); marks them expanded.
__device__ __forceinline__ int select_group(uint64_t *pool, int np, int mc, int &hint, uint32_t *sel)
{
    const int lane = lane_id();
    np = uni(np);
    mc = uni(mc);
    hint = uni(hint);
    int nsel = 0;
    for (int base = hint & ~(WAVE - 1); base < np && nsel < mc; base += WAVE) {
        const int i = base + lane;
        const uint64_t key = i < np ? pool[i] : EMPTY_KEY;
        const bool unexp = i < np && i >= hint && !(key & 1ull);
        const uint64_t bal = __ballot(unexp);
        // the first `need` unexpanded entries of the chunk, in pool order
        const int need = mc - nsel;
        const int rank = mbcnt(bal);
        if (unexp && rank < need) {
            pool[i] = key | 1ull;
            sel[nsel + rank] = key_id(key);
        }
        nsel += min(need, __popcll(bal));
        // next unexpanded entry: the chunk's first one not taken, else past it
        const uint64_t rest = __ballot(unexp && rank >= need);
        hint = rest ? base + __ffsll((unsigned long long)rest) - 1 : min(np, base + WAVE);
    }
    wave_sync();
    return nsel;
}

// Merge keys[0..n) (EMPTY entries allowed) into the sorted pool, keeping the
// ef smallest; returns the new size; pool/pool2 are swapped; `hint` becomes
// min(hint, first position a new entry landed at).
__device__ __forceinline__ int merge_into_pool(uint64_t *&pool, uint64_t *&pool2, int np, int ef, uint64_t *keys,
                                               int n, int &hint)
{
    const int lane = lane_id();
    np = uni(np);
    ef = uni(ef);
    n = uni(n);
    const uint64_t thr = (np >= ef) ? pool[ef - 1] : EMPTY_KEY;
    int ns = 0;
    for (int b = 0; b < n; b += WAVE) {
        const int i = b + lane;
        const uint64_t key = i < n ? keys[i] : EMPTY_KEY;
        const bool s = key < thr;
        const uint64_t bal = __ballot(s);
        wave_sync();
        if (s) keys[ns + mbcnt(bal)] = key;
        ns += __popcll(bal);
    }
    wave_sync();
    if (ns == 0) return np;
    // Merge survivors in batches of up to 64 (merging batch by batch gives the
    // same ef best).  Survivor s_k is broadcast by readlane; one compare per
    // lane (pool entry < s_k) gives, through the ballot mask, s_k's rank in
    // the pool (written to lane k) and, added up over k, how many survivors
    // do not precede the pool entry.  Pool and survivor keys are distinct
    // (distinct ids), so no compare needs an equality case.
    int mind = ef;
    for (int b = 0; b < ns; b += WAVE) {
        const int cnt = uni(min(WAVE, ns - b));
        const uint64_t sv = lane < cnt ? keys[b + lane] : EMPTY_KEY;
        int lo = 0, rs = 0;
        {
            // first 64 pool entries, with the survivors' ranks among themselves
            const uint64_t pv = lane < np ? pool[lane] : EMPTY_KEY;
            int lo_t = 0, after = 0;
            for (int kk = 0; kk < cnt; kk++) {
                const uint64_t s = readlane_u64(sv, kk);
                const bool lt = pv < s;
                lo_t = writelane_i32(lo_t, __popcll(__ballot(lt)), kk);
                after += lt ? 1 : 0;
                rs += (s < sv) ? 1 : 0;
            }
            lo = lo_t;
            const int dest = lane + (cnt - after);
            if (lane < np && dest < ef) pool2[dest] = pv;
        }
        for (int t = 1; t * WAVE < np; t++) {
            const int i = t * WAVE + lane;
            const uint64_t pv = i < np ? pool[i] : EMPTY_KEY;
            int lo_t = 0, after = 0;
            for (int kk = 0; kk < cnt; kk++) {
                const uint64_t s = readlane_u64(sv, kk);
                const bool lt = pv < s;
                lo_t = writelane_i32(lo_t, __popcll(__ballot(lt)), kk);
                after += lt ? 1 : 0;
            }
            lo += lo_t;
            const int dest = i + (cnt - after);
            if (i < np && dest < ef) pool2[dest] = pv;
        }
        const int dest = lo + rs;
        if (lane < cnt && dest < ef) pool2[dest] = sv;
        // the batch's smallest survivor (rank 0 among them) lands first
        const uint64_t first = __ballot(lane < cnt && rs == 0);
        const int d0 = __builtin_amdgcn_readlane(lo, __ffsll((unsigned long long)first) - 1);
        mind = min(mind, d0);
        wave_sync();
        uint64_t *t = pool;
        pool = pool2;
        pool2 = t;
        np = min(ef, np + cnt);
    }
    hint = min(hint, mind);
    return np;
}

template <int MROWS>
__device__ __forceinline__ void load_adj(const KParams &P, const uint32_t *sel, int nsel, int rpc, int32_t (&ids)[MROWS])
{
    const int lane = lane_id();
    const int nrows = nsel * rpc;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        ids[r] = -1;
        if (r < nrows) {
            const int ci = r / rpc, part = r - ci * rpc;
            const int j = part * WAVE + lane;
            if (j < P.r0) ids[r] = P.adj0[(size_t)sel[ci] * P.r0 + j];
        }
    }
}

// Visit the neighbour ids of one group, evaluate the new ones; results in
// L.lkeys[0..return).  Exactly one lane evaluates each id (the CAS winner).
#ifdef GVS_PROFILE
#define EXP_PROF , uint64_t (&prof_acc)[8], uint64_t &prof_t
#define EXP_PROF_ARGS , prof_acc, prof_t
#else
#define EXP_PROF
#define EXP_PROF_ARGS
#endif

// `frozen` (may be null) is the LDS hash of a query that outgrew it: its ids
// are visited and are looked up there first (read-only); the HBM table holds
// the ids visited since.  With fewer than 64 Bloom words every id the frozen
// hash does not hold is gathered speculatively.
template <int NCH, int STORE, int METRIC, int MROWS>
__device__ __forceinline__ int expand_group(const KParams &P, const WaveLds &L, const float *qf, uint32_t *T,
                                            const int32_t (&ids)[MROWS], int nrows, int &evals,
                                            const uint32_t *frozen EXP_PROF)
{
    const int lane = lane_id();
    const uint32_t nbw = L.nbw;
    const bool use_bloom = nbw >= 64u;
    const uint32_t tmask = (1u << P.vlog2) - 1u;
    const uint32_t hmask = (1u << P.hlog2) - 1u;
    uint32_t prev[MROWS], pos[MROWS];
    int32_t idv[MROWS];  // the row's ids, -1 where the frozen hash holds them
    int spos[MROWS];
    uint32_t spec_bits = 0u;
    int nl = 0;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        prev[r] = 0u;
        pos[r] = 0u;
        spos[r] = 0;
        if (r < nrows) {
            const int32_t id = ids[r];
            bool valid = id >= 0;
            if (frozen && P.hnarrow) {
                const uint16_t *F = (const uint16_t *)frozen;
                uint32_t hp = 0u, q = 0u;
                if (valid) nkey((uint32_t)id, P.hbits, P.hlog2, hp, q);
                bool probing = valid;
                uint32_t step = 0u;  // every probing lane is at the same step
                while (__ballot(probing)) {
                    if (probing) {
                        const uint32_t v = F[hp];
                        if (v == nentry(q, step)) {
                            valid = false;
                            probing = false;
                        } else if (v == 0u || step == 127u) {
                            probing = false;
                        } else {
                            hp = (hp + step + 1u) & hmask;
                        }
                    }
                    step++;
                }
            } else if (frozen) {
                const uint32_t want = (uint32_t)id + 1u;
                uint32_t hp = valid ? vhash((uint32_t)id, P.hlog2) : 0u;
                bool probing = valid;
                uint32_t step = 0u;  // every probing lane is at the same step
                while (__ballot(probing)) {
                    step++;
                    if (probing) {
                        const uint32_t v = frozen[hp];
                        if (v == want) {
                            valid = false;
                            probing = false;
                        } else if (v == 0u) {
                            probing = false;
                        } else {
                            hp = (hp + step) & hmask;
                        }
                    }
                }
            }
            idv[r] = valid ? id : -1;
            bool maybe = false;
            if (valid) {
                if (use_bloom) {
                    uint32_t w;
                    uint64_t m;
                    bloom_pos((uint32_t)id, P.bloom_seed, nbw, w, m);
                    const uint64_t old = atomicOr((unsigned long long *)&L.bloom[w], (unsigned long long)m);
                    maybe = (old & m) == m;
                }
                pos[r] = vhash((uint32_t)id, P.vlog2);
                prev[r] = atomicCAS(&T[pos[r]], 0u, (uint32_t)id + 1u);
            }
            const bool spec = valid && !maybe;
            const uint64_t bal = __ballot(spec);
            if (spec) {
                spos[r] = nl + mbcnt(bal);
                L.lids[spos[r]] = (uint32_t)id;
            }
            nl += __popcll(bal);
            spec_bits |= (spec ? 1u : 0u) << r;
        }
    }
    wave_sync();
    PROF_MARK(2);
    // ids the Bloom filter has never seen: gather while the CAS is in flight
    eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, nl, L.lkeys);
    PROF_MARK(3);
    int nl2 = nl;
    // table collisions of all rows resolved together, one slot per round
    // (as in expand_lds: an id met in two rows moves in step and is won once)
    uint32_t pend = 0u;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        const uint32_t want = (uint32_t)idv[r] + 1u;
        if (r < nrows && idv[r] >= 0 && prev[r] != 0u && prev[r] != want) pend |= 1u << r;
    }
    for (uint32_t round = 1u; __ballot(pend != 0u); round++) {
#pragma unroll
        for (int r = 0; r < MROWS; r++) {
            if (pend & (1u << r)) {
                const uint32_t want = (uint32_t)idv[r] + 1u;
                pos[r] = (pos[r] + round) & tmask;
                prev[r] = atomicCAS(&T[pos[r]], 0u, want);
                if (prev[r] == 0u || prev[r] == want) pend &= ~(1u << r);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        if (r < nrows) {
            const int32_t id = idv[r];
            const bool valid = id >= 0;
            const bool won = valid && prev[r] == 0u;
            const bool spec = (spec_bits >> r) & 1u;
            if (spec && !won) L.lkeys[spos[r]] = EMPTY_KEY;  // same id twice in one group
            const bool late = won && !spec;                   // Bloom false positive
            const uint64_t bal = __ballot(late);
            if (late) L.lids[nl2 + mbcnt(bal)] = (uint32_t)id;
            nl2 += __popcll(bal);
            evals += __popcll(__ballot(won));
        }
    }
    wave_sync();
    PROF_MARK(4);
    if (nl2 > nl) {
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids + nl, nl2 - nl, L.lkeys + nl);
        wave_sync();
    }
    PROF_MARK(5);
    return nl2;
}

// ---- exact visited set in LDS --------------------------------------------
// Open addressing over 2^hlog2 slots of this wave's LDS (key = id + 1), so a
// lookup-and-insert is one LDS compare-and-swap and collisions cost LDS
// latency only.  Every table (LDS hash, HBM table) probes triangularly:
// step d looks at home + d(d+1)/2, which reaches every slot of a
// power-of-two table and avoids the long clusters of linear probing near
// 3/4 load; all lanes probing together are at the same step.  LDS CAS
// serialises lanes, so an id met twice in one group is evaluated once.
// Used while the visited count stays below 3/4 of the slots; a query that
// outgrows it freezes it and goes on with the HBM table.

// The same with 2-byte slots (ef above 128).  A lane's compare-and-swap guesses that the
// slot's word is empty; a failed one returns the word, so a later step in
// the same word compares against what is there (a lane that retries a slot
// or steps within its word issues no atomic that round).  Returns -1 if a
// probe sequence ran past step 127 (the query overflows).
template <int MROWS>
__device__ __forceinline__ int visit_lds_narrow(const KParams &P, const WaveLds &L, const int32_t (&ids)[MROWS],
                                                int nrows)
{
    const uint32_t hmask = (1u << P.hlog2) - 1u;
    uint32_t sl[MROWS], cur[MROWS], st[MROWS];
    uint32_t pend = 0u, won = 0u;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        sl[r] = 0u;
        cur[r] = 0u;
        st[r] = 0u;
        if (r < nrows && ids[r] >= 0) {
            uint32_t q;
            nkey((uint32_t)ids[r], P.hbits, P.hlog2, sl[r], q);
            cur[r] = atomicCAS(&L.vh[sl[r] >> 1], 0u, nentry(q, 0u) << ((sl[r] & 1u) << 4));
            if (cur[r] == 0u) won |= 1u << r;
            else pend |= 1u << r;
        }
    }
    bool fail = false;
    while (__ballot(pend != 0u)) {
#pragma unroll
        for (int r = 0; r < MROWS; r++) {
            if (pend & (1u << r)) {
                uint32_t home, q;
                nkey((uint32_t)ids[r], P.hbits, P.hlog2, home, q);
                const uint32_t sh = (sl[r] & 1u) << 4;
                const uint32_t h = (cur[r] >> sh) & 0xFFFFu;
                const uint32_t want = nentry(q, st[r]);
                if (h == want) {
                    pend &= ~(1u << r);  // visited
                } else if (h != 0u) {
                    if (st[r] == 127u) {
                        fail = true;
                        pend &= ~(1u << r);
                    } else {
                        st[r]++;
                        const uint32_t ns = (sl[r] + st[r]) & hmask;
                        if ((ns >> 1) != (sl[r] >> 1)) cur[r] = 0u;
                        sl[r] = ns;
                    }
                } else {
                    const uint32_t old = atomicCAS(&L.vh[sl[r] >> 1], cur[r], cur[r] | (want << sh));
                    if (old == cur[r]) {
                        won |= 1u << r;
                        pend &= ~(1u << r);
                    } else {
                        cur[r] = old;
                    }
                }
            }
        }
    }
    int nl = 0;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        if (r < nrows) {
            const bool w = (won >> r) & 1u;
            const uint64_t bal = __ballot(w);
            if (w) L.lids[nl + mbcnt(bal)] = (uint32_t)ids[r];
            nl += __popcll(bal);
        }
    }
    return __ballot(fail) ? -1 : nl;
}

template <int NCH, int STORE, int METRIC, int MROWS>
__device__ __forceinline__ int expand_lds(const KParams &P, const WaveLds &L, const float *qf,
                                          const int32_t (&ids)[MROWS], int nrows, int &evals EXP_PROF)
{
    if (P.hnarrow) {
        const int nl = visit_lds_narrow<MROWS>(P, L, ids, nrows);
        if (nl < 0) return -1;
        evals += nl;
        wave_sync();
        PROF_MARK(2);
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, nl, L.lkeys);
        wave_sync();
        PROF_MARK(3);
        return nl;
    }
    const uint32_t hmask = (1u << P.hlog2) - 1u;
    // First probe of every row, issued back to back so the LDS atomics of
    // the rows overlap.  Collisions are then resolved for all rows together:
    // each round moves every pending (lane, row) one slot on, so the wave
    // waits for the longest probe sequence of the group, not for the sum of
    // each row's longest.  Within a round the rows probe in row order, and
    // an id met in two rows starts at the same slot in the same round and
    // moves in step, so it is won once.
    uint32_t pos[MROWS], prev[MROWS];
    uint32_t pend = 0u;  // bit r: row r still probing
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        pos[r] = 0u;
        prev[r] = ~0u;
        if (r < nrows && ids[r] >= 0) {
            pos[r] = vhash((uint32_t)ids[r], P.hlog2);
            prev[r] = atomicCAS(&L.vh[pos[r]], 0u, (uint32_t)ids[r] + 1u);
        }
    }
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        const uint32_t want = (uint32_t)ids[r] + 1u;
        if (r < nrows && ids[r] >= 0 && prev[r] != 0u && prev[r] != want) pend |= 1u << r;
    }
    for (uint32_t round = 1u; __ballot(pend != 0u); round++) {
#pragma unroll
        for (int r = 0; r < MROWS; r++) {
            if (pend & (1u << r)) {
                const uint32_t want = (uint32_t)ids[r] + 1u;
                pos[r] = (pos[r] + round) & hmask;
                prev[r] = atomicCAS(&L.vh[pos[r]], 0u, want);
                if (prev[r] == 0u || prev[r] == want) pend &= ~(1u << r);
            }
        }
    }
    int nl = 0;
#pragma unroll
    for (int r = 0; r < MROWS; r++) {
        if (r < nrows) {
            const int32_t id = ids[r];
            const bool won = id >= 0 && prev[r] == 0u;
            const uint64_t bal = __ballot(won);
            if (won) L.lids[nl + mbcnt(bal)] = (uint32_t)id;
            nl += __popcll(bal);
        }
    }
    evals += nl;
    wave_sync();
    PROF_MARK(2);
    eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, nl, L.lkeys);
    wave_sync();
    PROF_MARK(3);
    return nl;
}

// Layer-0 entry point and its distance: the given entry, or hnswlib's
// upper-layer greedy descent (HNSWLIB/hnswalg.h:1119-1147): from the top
// level down, move to the closest neighbour while it is strictly closer.
// Also returns the nodes the descent settled on at levels 3 and 2 (the entry
// when the graph has fewer levels): a coarse address of the query's region.
template <int NCH, int STORE, int METRIC>
__device__ __forceinline__ void find_entry_path(const KParams &P, int64_t qi, const WaveLds &L, const float *qf,
                                                uint32_t &ep, float &epd, int &up_hops, int &up_evals,
                                                uint32_t &n2, uint32_t &n3)
{
    const int lane = lane_id();
    const int32_t given = P.entries ? P.entries[qi] : -1;  // a negative entry: descend as without one
    if (given >= 0) {
        ep = (uint32_t)given;
        if (lane == 0) L.lids[0] = ep;
        wave_sync();
        eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, 1, L.lkeys);
        wave_sync();
        epd = key_dist(L.lkeys[0]);
        n2 = n3 = ep;
        return;
    }
    uint32_t cur = (uint32_t)P.entry;
    n2 = n3 = cur;
    if (lane == 0) L.lids[0] = cur;
    wave_sync();
    eval_rows<NCH, STORE, METRIC>(P, qf, L.lids, 1, L.lkeys);
    wave_sync();
    float curd = key_dist(L.lkeys[0]);
    up_evals = 1;
    for (int level = P.max_level; level > 0; level--) {
        for (;;) {
            const int64_t row = P.upper_off[cur] + (level - 1);
            const int32_t nid = lane < P.ru ? P.adj_up[row * P.ru + lane] : -1;
            const uint64_t bal = __ballot(nid >= 0);
            const int n = __popcll(bal);
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
                const float bd = unord_f((uint32_t)(best >> 32));
                if (bd < curd) {  // strictly closer
                    curd = bd;
                    cur = L.lids[(int)(best & 63u)];
                    wave_sync();
                    continue;
                }
            }
            break;
        }
        n3 = level == 3 ? cur : n3;
        n2 = level == 2 ? cur : n2;
    }
    ep = cur;
    epd = curd;
    if (P.max_level < 2) n2 = cur;
    if (P.max_level < 3) n3 = n2;
}

template <int NCH, int STORE, int METRIC>
__device__ __forceinline__ void find_entry(const KParams &P, int64_t qi, const WaveLds &L, const float *qf,
                                           uint32_t &ep, float &epd, int &up_hops, int &up_evals)
{
    uint32_t n2, n3;
    find_entry_path<NCH, STORE, METRIC>(P, qi, L, qf, ep, epd, up_hops, up_evals, n2, n3);
}

// The query in the lane layout eval_rows reads (lane p = lane & 7 holds the
// elements (p + 8 i) * E + t), zero past dim.  A query in host memory
// (sys_query: system-scope loads, fresh across the bus) is read once: lane l
// loads element 64 b + l of each 64-element block b (one contiguous 256-byte
// read per block), and the layout is built with cross-lane permutes, instead
// of every lane reading its own elements across the bus (8 lanes share each).
template <int NCH, int E>
__device__ __forceinline__ void load_query(const KParams &P, const float *Q, bool sys_query, float *qf)
{
    const int lane = lane_id();
    const int p = lane & 7;
    if constexpr (E == 4) {
        if (sys_query) {
            // element 32 i + 4 p + t lies in block i >> 1, at 32 (i & 1) + 4 p + t
            constexpr int NB = (NCH + 1) / 2;
            float blk[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int idx = 64 * b + lane;
                blk[b] = idx < P.dim ? __hip_atomic_load(&Q[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < NCH; i++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int src = 32 * (i & 1) + 4 * p + t;
                    qf[i * 4 + t] = __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(blk[i >> 1])));
                }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < NCH; i++)
#pragma unroll
        for (int t = 0; t < E; t++) {
            const int idx = (p + 8 * i) * E + t;
            float v = 0.0f;
            if (idx < P.dim)
                v = sys_query ? __hip_atomic_load(&Q[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : Q[idx];
            qf[i * E + t] = v;
        }
}

// One result row: plain stores, or, for a request whose results cross the
// bus (sys), system-scope stores that write through to host memory, so the
// completion flag needs only a wait for them, not a write-back of L2.
__device__ __forceinline__ void put_result(const KParams &P, int64_t at, int64_t lab, float d, bool sys)
{
    if (sys) {
        __hip_atomic_store(&P.out_labels[at], lab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&P.out_dists[at], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        P.out_labels[at] = lab;
        P.out_dists[at] = d;
    }
}

__device__ __forceinline__ void put_stat(int32_t *row, int i, int32_t v, bool sys)
{
    if (sys)
        __hip_atomic_store(&row[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        row[i] = v;
}

// Returns true if the query's visited table overflowed (its results are
// then label -1 / +inf and its status word is 1).
// POOL: the query may start from an initial pool (init, init_len) instead
// of the descent; a compile-time switch, so that the other kernels hold none
// of that code.
template <int NCH, int STORE, int METRIC, int MROWS, bool POOL = false>
__device__ bool search_query_req(const KParams &P, int64_t qi, int64_t slot, unsigned char *lds, int k, int ef,
                                 const float *Q, bool sys_query, const int32_t *init = nullptr, int init_len = 0)
{
    if constexpr (!POOL) {
        init = nullptr;
        init_len = 0;
    }
    constexpr int E = StoreTraits<STORE>::E;
    const int lane = lane_id();
    const int p = lane & 7;
    WaveLds L = carve(P, lds);

    PROF_DECL
    if (ef < k) ef = k;

    float qf[NCH * E];
    load_query<NCH, E>(P, Q, sys_query, qf);

    int up_hops = 0, up_evals = 0;
    uint32_t ep = 0u;
    float epd = 0.0f;
    if (!init) find_entry<NCH, STORE, METRIC>(P, qi, L, qf, ep, epd, up_hops, up_evals);

    PROF_MARK(0);
    // layer-0 search, HNSWLIB/hnswalg.h:240-326; groups of up to mc
    // candidates, up to mg groups in flight (INTER/scheduler.hpp:100-157).
    // A group is selected from the pool as it stands after the oldest
    // in-flight group has been merged; the next group's adjacency loads are
    // issued before the current group is gathered, so with mg = 2 they
    // overlap.
    uint64_t *pool = L.poolA, *pool2 = L.poolB;
    int np = 1, hint = 0;
    int hops = 0, evals = 1;
    bool overflow = false;
    uint32_t *T = nullptr;
    const int rpc = (P.r0 + WAVE - 1) / WAVE;
    // Visited set: the LDS hash while it stays below 3/4 full (small ef);
    // a query that outgrows it freezes it and goes on with the HBM table for
    // the ids it meets from then on (frozen hash first, then the table).
    // Large ef does the same with the 2-byte LDS hash where the graph's ids
    // fit it, and otherwise starts with the HBM table and the LDS Bloom filter.
    bool lds_mode = P.hlog2 > 0;
    const uint32_t *frozen = nullptr;
    if (!lds_mode) T = P.vtab + ((size_t)slot << P.vlog2);
    int tlimit = lds_mode ? (int)((3u << P.hlog2) >> 2) - P.lcap : (int)((3u << P.vlog2) >> 2) - P.lcap;
    if (lds_mode) {
        uint4 *H4 = (uint4 *)L.vh;
        const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t n4 = (1u << P.hlog2) >> (P.hnarrow ? 3 : 2);
        for (uint32_t i = lane; i < n4; i += WAVE) H4[i] = z4;
    } else {
        for (uint32_t w = lane; w < L.nbw; w += WAVE) L.bloom[w] = 0ull;
    }
    wave_sync();
    if (init) {
        np = 0;
        evals = 0;
    } else if (lane == 0) {
        pool[0] = mk_key(epd, ep);
        if (lds_mode && P.hnarrow) {
            uint32_t home, q;
            nkey(ep, P.hbits, P.hlog2, home, q);
            L.vh[home >> 1] = nentry(q, 0u) << ((home & 1u) << 4);
        } else if (lds_mode) {
            L.vh[vhash(ep, P.hlog2)] = ep + 1u;
        } else {
            uint32_t w;
            uint64_t m;
            bloom_pos(ep, P.bloom_seed, L.nbw, w, m);
            L.bloom[w] |= m;
            atomicCAS(&T[vhash(ep, P.vlog2)], 0u, ep + 1u);
        }
    }
    wave_sync();

    int nres = 0;
    bool haveQ = false, haveGroup = false;
    // DST pipeline: up to mg groups in flight, i.e. selected and not yet
    // merged.  Selected groups wait in a ring of mg lists of up to 8 ids in
    // LDS (list sizes 4 bits per ring position in gsz); each step expands the
    // oldest, after issuing the adjacency loads of the one behind it, so
    // those overlap its gathers.  A group is selected from the pool as it
    // stands once the oldest in-flight group has been merged.
    // An initial pool (the NSG search's start, nsg.py:66-90) goes first, in
    // chunks of MROWS rows of ids that take a group's place in the same
    // expansion (one copy of it in the kernel) and are merged like groups.
    const int mg = P.mg;
    uint32_t gsz = 0u;
    int gh = 0, pend = 0, inflight = 0;
    int ic = init ? 0 : init_len;  // next initial-pool id
    bool pLoaded = false;
    int32_t idsP[MROWS], idsN[MROWS];
    for (;;) {
        if (haveQ) {  // complete the oldest group (or pool chunk)
            np = merge_into_pool(pool, pool2, np, ef, L.lkeys, nres, hint);
            haveQ = false;
            if (haveGroup) inflight--;
        }
        PROF_MARK(6);
        int nrowsP;
        int gn = gh;
        if (POOL && ic < init_len) {
#pragma unroll
            for (int r = 0; r < MROWS; r++) {
                const int j = ic + r * WAVE + lane;
                const int32_t id = j < init_len ? init[j] : -1;
                idsP[r] = (id >= 0 && (int64_t)id < P.n) ? id : -1;
            }
            nrowsP = min(MROWS, (init_len - ic + WAVE - 1) / WAVE);
            ic += MROWS * WAVE;
            haveGroup = false;
        } else {
            // refill the pipeline up to mg groups
            while (inflight < mg) {
                int pos = gh + pend;
                if (pos >= mg) pos -= mg;
                const int ns = select_group(pool, np, P.mc, hint, L.sel + 8 * pos);
                if (ns == 0) break;
                gsz = (gsz & ~(15u << (4 * pos))) | ((uint32_t)ns << (4 * pos));
                pend++;
                inflight++;
                hops += ns;
            }
            if (pend == 0) break;
            const int nselP = (int)((gsz >> (4 * gh)) & 15u);
            if (!pLoaded) load_adj<MROWS>(P, L.sel + 8 * gh, nselP, rpc, idsP);
            gn = gh + 1 == mg ? 0 : gh + 1;
            pLoaded = pend > 1;
            if (pLoaded) load_adj<MROWS>(P, L.sel + 8 * gn, (int)((gsz >> (4 * gn)) & 15u), rpc, idsN);
            nrowsP = nselP * rpc;
            haveGroup = true;
        }
        PROF_MARK(1);
        if (lds_mode)
            nres = expand_lds<NCH, STORE, METRIC, MROWS>(P, L, qf, idsP, nrowsP, evals EXP_PROF_ARGS);
        else
            nres = expand_group<NCH, STORE, METRIC, MROWS>(P, L, qf, T, idsP, nrowsP, evals,
                                                           frozen EXP_PROF_ARGS);
        if (nres < 0) {  // probe sequence of the 2-byte hash too long
            overflow = true;
            break;
        }
        haveQ = true;
        if (haveGroup) {  // the next group moves up
            pend--;
            gh = gn;
#pragma unroll
            for (int r = 0; r < MROWS; r++) idsP[r] = idsN[r];
        }
        if (evals > tlimit) {
            if (lds_mode && P.vtab) {
                lds_mode = false;
                frozen = L.vh;
                T = P.vtab + ((size_t)slot << P.vlog2);
                tlimit = evals + (int)((3u << P.vlog2) >> 2) - P.lcap;
                for (uint32_t w = lane; w < L.nbw; w += WAVE) L.bloom[w] = 0ull;
                if (lane == 0) atomicAdd(&P.counters[2], 1u);
                wave_sync();
            } else {
                overflow = true;
                break;
            }
        }
    }

    // Low-precision traversal: rerank the final pool with exact fp32
    // distances (fp32 copy, the kernels' fp32 summation order), then sort by
    // (dist, id).
    if constexpr (STORE == STORE_BF16 || STORE == STORE_FP8) {
        if (P.rerank && !overflow) {
            constexpr int NCH32 = NCH * E / 4;
            float q32[NCH32 * 4];
#pragma unroll
            for (int i = 0; i < NCH32; i++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int idx = (p + 8 * i) * 4 + t;
                    float v = 0.0f;
                    if (idx < P.dim)
                        v = sys_query ? __hip_atomic_load(&Q[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : Q[idx];
                    q32[i * 4 + t] = v;
                }
            KParams P32 = P;
            P32.vec = P.vec_f32;
            P32.row_elems = P.row_elems_f32;
            for (int b = 0; b < np; b += P.lcap) {
                const int cnt = min(P.lcap, np - b);
                for (int i = lane; i < cnt; i += WAVE) L.lids[i] = key_id(pool[b + i]);
                wave_sync();
                eval_rows<NCH32, STORE_F32, METRIC>(P32, q32, L.lids, cnt, L.lkeys);
                wave_sync();
                for (int i = lane; i < cnt; i += WAVE) pool2[b + i] = L.lkeys[i];
                wave_sync();
            }
            // rank sort of the reranked keys
            for (int i = lane; i < np; i += WAVE) {
                const uint64_t v = pool2[i];
                int r = 0;
                for (int j = 0; j < np; j++) r += (pool2[j] < v) ? 1 : 0;
                pool[r] = v;
            }
            wave_sync();
        }
    }

    // results, closer first (HNSWLIB/hnswlib.h:171-185)
    const int nout = P.out_k_only ? min(k, P.k) : P.k;
    for (int i = lane; i < nout; i += WAVE) {
        const uint64_t key = (!overflow && i < np && i < k) ? pool[i] : EMPTY_KEY;
        int64_t lab = -1;
        float d = __builtin_inff();
        if (key != EMPTY_KEY) {
            const uint32_t id = key_id(key);
            lab = P.labels ? P.labels[id] : (int64_t)id;
            d = key_dist(key);
        }
        put_result(P, qi * P.k + i, lab, d, sys_query);
    }
    PROF_MARK(7);
    if (lane == 0) {
        if (P.out_stats) {
            int32_t *row = P.out_stats + qi * GVS_STATS_ROW;
            put_stat(row, 0, hops, sys_query);
            put_stat(row, 1, evals, sys_query);
            if (!P.keep_upper_stats) {
                put_stat(row, 2, up_hops, sys_query);
                put_stat(row, 3, up_evals, sys_query);
            }
#ifdef GVS_PROFILE
            for (int ph = 0; ph < 8; ph++) row[4 + ph] = (int32_t)(prof_acc[ph] >> 6);
#endif
        }
        if (P.status) P.status[qi] = overflow ? 1 : 0;
        if (overflow) atomicAdd(&P.counters[1], 1u);
    }
    // leave the slot's visited table empty for the next query
    if (T) {
        uint4 *T4 = (uint4 *)T;
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t i = lane; i < ((1u << P.vlog2) >> 2); i += WAVE) T4[i] = z;
        // only this wave uses the slot's table, and only through L2 atomics:
        // the cleared lines must have reached L2 before its next CAS
        __builtin_amdgcn_s_waitcnt(0);
    }
    wave_sync();
    return overflow;
}

template <int NCH, int STORE, int METRIC, int MROWS, bool POOL = false>
__device__ __forceinline__ void search_query(const KParams &P, int64_t qi, int64_t slot, unsigned char *lds)
{
    const int64_t qrow = P.q_index ? P.q_index[qi] : qi;
    const int k = P.q_k ? P.q_k[qi] : P.k;
    const int ef = P.q_ef ? P.q_ef[qi] : P.ef;
    const int32_t *init = nullptr;
    int init_len = 0;
    if constexpr (POOL) {
        init = P.init_mode == 1 ? P.init_ids + qrow * P.init_ld : P.adj0 + (size_t)P.entry * P.r0;
        init_len = P.init_mode == 1 ? P.init_len : min(P.r0, max(ef, k));
    }
    search_query_req<NCH, STORE, METRIC, MROWS, POOL>(P, qi, slot, lds, k, ef, P.queries + qrow * P.ldq, false,
                                                      init, init_len);
}

// ---- several waves per query (intra-query parallelism) -------------------
// One workgroup of W waves owns one query (the reference's INTRA design
// spreads one query's neighbour work over channels, INTRA/utils.hpp:396-570).
// Each step: wave 0 merges the lists of the group expanded last and refills
// the pipeline of selected groups (up to mg in flight, each of up to mc
// candidates); the oldest group's mc * ceil(r0/64) adjacency rows are dealt
// round-robin to the W waves, which issue the loads of their rows of the
// next group, then test the current rows against the shared visited set and
// gather and evaluate the new ids into their own key lists.  The visited set
// is exact and shared, groups are expanded in selection order, and the merge
// keeps the ef smallest keys whatever the order of the lists, so the result,
// hops and evaluation counts equal those of one wave running DST(mc, mg).
__device__ __forceinline__ void wg_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup LDS: shared part, then WROWS*64 keys and ids per wave.
struct WgLds {
    uint64_t *poolA, *poolB, *bloom, *lkeys;
    uint32_t *lids, *sel, *ctl, *vh;
};

template <int WROWS>
__device__ __forceinline__ WgLds carve_wg(const KParams &P, unsigned char *base, int W)
{
    WgLds G;
    G.poolA = (uint64_t *)base;
    G.poolB = G.poolA + P.ef_cap;
    G.bloom = G.poolB + P.ef_cap;
    G.lkeys = G.bloom + (1u << P.blog2);
    G.lids = (uint32_t *)(G.lkeys + W * WROWS * WAVE);
    G.sel = G.lids + W * WROWS * WAVE;
    G.ctl = G.sel + 8 * P.mg;
    G.vh = G.ctl + WG_CTL_WORDS;
    return G;
}

// ctl words: the group to expand (size, ring position), the group behind it
// (size, 0 if none, and position), evaluations so far, visited-mode switch,
// overflow stop, the smallest pool place a new key took in the last merge,
// then per wave the number of keys it offers to the merge
constexpr int WG_NSEL = 0, WG_EVALS = 1, WG_MODE = 2, WG_STOP = 3, WG_POS = 4, WG_NNEXT = 5, WG_PNEXT = 6,
              WG_HINT = 7, WG_CNT = 8;

// Merge by all W waves of a workgroup: list j (lists + j * lstride) holds
// cnt[j] keys, unsorted, each below the pool's ef-th key; keys are distinct
// (distinct ids).  The place of a key among the ef best is the number of
// pool keys below it plus the number of list keys below it, so wave w
// places its own list's keys (the pool rank by binary search, the list rank
// by counting over every list) and the pool entries (t * W + w) * 64 + lane.
// Writes pool2 (merged: the caller swaps the pools), returns the new pool
// size, and lowers *hint_word to the smallest place a list key took.
__device__ __forceinline__ int wg_merge(const uint64_t *pool, uint64_t *pool2, int np, int ef,
                                        const uint64_t *lists, int lstride, const uint32_t *cnt, int W, int w,
                                        uint32_t *hint_word, bool &merged)
{
    const int lane = lane_id();
    int S = 0;
    for (int j = 0; j < W; j++) S += (int)cnt[j];
    S = uni(S);
    merged = S > 0;
    if (S == 0) return np;
    const int cw = uni((int)cnt[w]);
    const uint64_t *mine = lists + (size_t)w * lstride;
    const int npw = np > w * WAVE ? (np - w * WAVE + W * WAVE - 1) / (W * WAVE) : 0;
    const int passes = max(npw, (cw + WAVE - 1) / WAVE);
    int top = 1;  // largest power of two <= np
    while (2 * top <= np) top *= 2;
    int mind = ef;
    for (int t = 0; t < passes; t++) {
        const int si = t * WAVE + lane;
        const bool hs = si < cw;
        const uint64_t sv = hs ? mine[si] : EMPTY_KEY;
        const int pi = (t * W + w) * WAVE + lane;
        const bool hp = pi < np;
        const uint64_t pv = hp ? pool[pi] : EMPTY_KEY;
        int cs = 0, cp = 0;
        for (int j = 0; j < W; j++) {
            const uint64_t *lj = lists + (size_t)j * lstride;
            const int cj = uni((int)cnt[j]);
#pragma unroll 8
            for (int m = 0; m < cj; m++) {
                const uint64_t v = lj[m];
                cs += v < sv ? 1 : 0;
                cp += v < pv ? 1 : 0;
            }
        }
        if (hp && pi + cp < ef) pool2[pi + cp] = pv;
        if (hs) {
            int lo = 0;  // pool keys below sv
            for (int step = top; step > 0; step >>= 1)
                if (lo + step <= np && pool[lo + step - 1] < sv) lo += step;
            const int d = lo + cs;
            if (d < ef) pool2[d] = sv;
            mind = min(mind, d);
        }
    }
    const uint32_t m = (uint32_t)wave_min_u64((uint64_t)(uint32_t)mind);
    if (lane == 0) atomicMin(hint_word, m);
    return min(ef, np + S);
}

// One query with its own k and ef (query Q; sys_query: Q is in host memory,
// read with system-scope loads).  Returns true on every wave if the
// query's visited table overflowed.
template <int NCH, int STORE, int METRIC, int WROWS>
__device__ bool search_query_mw_req(const KParams &P, int64_t qi, int64_t slot, unsigned char *lds, int k, int ef,
                                    const float *Q, bool sys_query)
{
    constexpr int E = StoreTraits<STORE>::E;
    const int lane = lane_id();
    const int p = lane & 7;
    const int W = P.waves_per_block;
    const int w = threadIdx.x >> 6;
    WgLds G = carve_wg<WROWS>(P, lds, W);
    PROF_DECL
    // this wave's view: its own key/id lists, the shared visited structures
    WaveLds L;
    L.poolA = G.poolA;
    L.poolB = G.poolB;
    L.lkeys = G.lkeys + w * WROWS * WAVE;
    L.lids = G.lids + w * WROWS * WAVE;
    L.bloom = G.bloom;
    L.vh = G.vh;
    L.sel = G.sel;
    L.nbw = 1u << P.blog2;

    if (ef < k) ef = k;
    float qf[NCH * E];
    load_query<NCH, E>(P, Q, sys_query, qf);
    // every wave runs the (cheap, deterministic) descent on its own lists
    int up_hops = 0, up_evals = 0;
    uint32_t ep;
    float epd;
    find_entry<NCH, STORE, METRIC>(P, qi, L, qf, ep, epd, up_hops, up_evals);
    PROF_MARK(0);

    bool lds_mode = P.hlog2 > 0;
    const uint32_t *frozen = nullptr;
    uint32_t *T = lds_mode ? nullptr : P.vtab + ((size_t)slot << P.vlog2);
    int tlimit = lds_mode ? (int)((3u << P.hlog2) >> 2) - P.lcap : (int)((3u << P.vlog2) >> 2) - P.lcap;
    const int tid = threadIdx.x, nthr = W * WAVE;
    if (lds_mode) {
        for (uint32_t i = tid; i < (1u << P.hlog2); i += nthr) G.vh[i] = 0u;
    } else {
        for (uint32_t i = tid; i < L.nbw; i += nthr) G.bloom[i] = 0ull;
    }
    if (tid < WG_CTL_WORDS) G.ctl[tid] = 0u;
    wg_sync();
    uint64_t *pool = G.poolA, *pool2 = G.poolB;
    int np = 1, hint = 0, hops = 0;
    if (w == 0 && lane == 0) {
        pool[0] = mk_key(epd, ep);
        G.ctl[WG_EVALS] = 1u;
        if (lds_mode) {
            G.vh[vhash(ep, P.hlog2)] = ep + 1u;
        } else {
            uint32_t bw;
            uint64_t m;
            bloom_pos(ep, P.bloom_seed, L.nbw, bw, m);
            G.bloom[bw] |= m;
            atomicCAS(&T[vhash(ep, P.vlog2)], 0u, ep + 1u);
        }
    }
    wg_sync();
    const int rpc = (P.r0 + WAVE - 1) / WAVE;
    bool overflow = false;
    // wave 0's pipeline state (DST groups, as in the one-wave search)
    const int mg = P.mg;
    uint32_t gsz = 0u;
    int gh = 0, pend = 0, inflight = 0;
    bool haveQ = false;
    // every wave: its rows of the current group, and of the next one when
    // their loads were issued during the previous step
    int32_t idsP[WROWS], idsN[WROWS];
    bool pLoaded = false;
#ifdef GVS_PROFILE
    int steps = 0;
#endif
    for (;;) {
        if (w == 0) {
            if (haveQ) {  // the oldest group was merged at the end of the last step
                haveQ = false;
                inflight--;
                hint = min(hint, (int)G.ctl[WG_HINT]);
                const int evals = (int)G.ctl[WG_EVALS];
                if (evals > tlimit) {
                    if (lds_mode && P.vtab) {
                        for (uint32_t i = lane; i < L.nbw; i += WAVE) G.bloom[i] = 0ull;
                        if (lane == 0) {
                            G.ctl[WG_MODE] = 1u;
                            atomicAdd(&P.counters[2], 1u);
                        }
                        tlimit = evals + (int)((3u << P.vlog2) >> 2) - P.lcap;
                    } else if (lane == 0) {
                        G.ctl[WG_STOP] = 1u;
                    }
                }
            }
            while (inflight < mg) {  // refill the pipeline up to mg groups
                int pos = gh + pend;
                if (pos >= mg) pos -= mg;
                const int ns = select_group(pool, np, P.mc, hint, G.sel + 8 * pos);
                if (ns == 0) break;
                gsz = (gsz & ~(15u << (4 * pos))) | ((uint32_t)ns << (4 * pos));
                pend++;
                inflight++;
                hops += ns;
            }
            const int gn = gh + 1 == mg ? 0 : gh + 1;
            if (lane == 0) {
                G.ctl[WG_NSEL] = pend > 0 ? (gsz >> (4 * gh)) & 15u : 0u;
                G.ctl[WG_POS] = (uint32_t)gh;
                G.ctl[WG_NNEXT] = pend > 1 ? (gsz >> (4 * gn)) & 15u : 0u;
                G.ctl[WG_PNEXT] = (uint32_t)gn;
                G.ctl[WG_HINT] = 0x7fffffffu;
            }
            if (pend > 0) {
                pend--;
                gh = gn;
                haveQ = true;
            }
        }
        PROF_MARK(7);
        wg_sync();
        PROF_MARK(1);
        if (G.ctl[WG_STOP]) {
            overflow = true;
            break;
        }
        if (lds_mode && G.ctl[WG_MODE]) {
            lds_mode = false;
            frozen = G.vh;
            T = P.vtab + ((size_t)slot << P.vlog2);
        }
        const int nsel = (int)G.ctl[WG_NSEL];
        if (nsel == 0) break;
#ifdef GVS_PROFILE
        steps++;
#endif
        const int nnext = (int)G.ctl[WG_NNEXT];
        const uint32_t *selC = G.sel + 8 * G.ctl[WG_POS];
        const uint32_t *selN = G.sel + 8 * G.ctl[WG_PNEXT];
        // this wave's rows r = w + W * j of a group of n candidates
        const auto rows_of = [&](const uint32_t *sl, int n, int32_t(&ids)[WROWS]) {
#pragma unroll
            for (int j = 0; j < WROWS; j++) {
                ids[j] = -1;
                const int r = w + W * j;
                if (r < n * rpc) {
                    const int ci = r / rpc, jj = (r - ci * rpc) * WAVE + lane;
                    if (jj < P.r0) ids[j] = P.adj0[(size_t)sl[ci] * P.r0 + jj];
                }
            }
        };
        if (!pLoaded) rows_of(selC, nsel, idsP);
        pLoaded = nnext > 0;
        if (pLoaded) rows_of(selN, nnext, idsN);
        const int nrows = nsel * rpc;
        const int myrows = w < nrows ? min(WROWS, (nrows - w + W - 1) / W) : 0;
        int evals_w = 0, nres = 0;
        if (myrows > 0) {
            if (lds_mode)
                nres = expand_lds<NCH, STORE, METRIC, WROWS>(P, L, qf, idsP, myrows, evals_w EXP_PROF_ARGS);
            else
                nres = expand_group<NCH, STORE, METRIC, WROWS>(P, L, qf, T, idsP, myrows, evals_w,
                                                               frozen EXP_PROF_ARGS);
        }
        // keep the keys that can enter the pool (below its ef-th key),
        // compacted in place
        {
            const uint64_t thr = np >= ef ? pool[ef - 1] : EMPTY_KEY;
            int c = 0;
            for (int b = 0; b < nres; b += WAVE) {
                const int i = b + lane;
                const uint64_t key = i < nres ? L.lkeys[i] : EMPTY_KEY;
                const bool sv = key < thr;
                const uint64_t bal = __ballot(sv);
                wave_sync();
                if (sv) L.lkeys[c + mbcnt(bal)] = key;
                c += __popcll(bal);
            }
            if (lane == 0) {
                G.ctl[WG_CNT + w] = (uint32_t)c;
                if (evals_w) atomicAdd(&G.ctl[WG_EVALS], (uint32_t)evals_w);
            }
        }
#pragma unroll
        for (int j = 0; j < WROWS; j++) idsP[j] = idsN[j];
        PROF_MARK(4);
        wg_sync();
        PROF_MARK(5);
        // every wave merges its share; the pools swap when keys came in
        bool merged = false;
        np = wg_merge(pool, pool2, np, ef, G.lkeys, WROWS * WAVE, G.ctl + WG_CNT, W, w, G.ctl + WG_HINT, merged);
        if (merged) {
            uint64_t *t = pool;
            pool = pool2;
            pool2 = t;
        }
        wg_sync();
        PROF_MARK(6);
    }

    PROF_MARK(7);
    if (w == 0) {
        const int nout = P.out_k_only ? min(k, P.k) : P.k;
        for (int i = lane; i < nout; i += WAVE) {
            const uint64_t key = (!overflow && i < np && i < k) ? pool[i] : EMPTY_KEY;
            int64_t lab = -1;
            float d = __builtin_inff();
            if (key != EMPTY_KEY) {
                const uint32_t id = key_id(key);
                lab = P.labels ? P.labels[id] : (int64_t)id;
                d = key_dist(key);
            }
            put_result(P, qi * P.k + i, lab, d, sys_query);
        }
        if (lane == 0) {
            if (P.out_stats) {
                int32_t *row = P.out_stats + qi * GVS_STATS_ROW;
                put_stat(row, 0, hops, sys_query);
                put_stat(row, 1, (int)G.ctl[WG_EVALS], sys_query);
                if (!P.keep_upper_stats) {  // region order: the descent kernel wrote them
                    put_stat(row, 2, up_hops, sys_query);
                    put_stat(row, 3, up_evals, sys_query);
                }
#ifdef GVS_PROFILE
                // wave 0's phases: descent, wait at the step barrier, visit,
                // gather + distances, filter, wait at the list barrier,
                // merge (with its barrier), select + output
                for (int ph = 0; ph < 8; ph++) row[4 + ph] = (int32_t)(prof_acc[ph] >> 6);
                row[12] = (int32_t)steps;
#endif
            }
            if (P.status) P.status[qi] = overflow ? 1 : 0;
            if (overflow) atomicAdd(&P.counters[1], 1u);
        }
    }
    // leave the query's visited table empty for the next query of this slot
    if (T) {
        uint4 *T4 = (uint4 *)T;
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t i = tid; i < ((1u << P.vlog2) >> 2); i += nthr) T4[i] = z;
        __builtin_amdgcn_s_waitcnt(0);
    }
    wg_sync();
    return overflow;
}

template <int NCH, int STORE, int METRIC, int WROWS>
__device__ __forceinline__ void search_query_mw(const KParams &P, int64_t qi, int64_t slot, unsigned char *lds)
{
    const int64_t qrow = P.q_index ? P.q_index[qi] : qi;  // a retry row re-runs query q_index[qi]
    search_query_mw_req<NCH, STORE, METRIC, WROWS>(P, qi, slot, lds, P.k, max(P.ef, P.k), P.queries + qrow * P.ldq,
                                                   false);
}

}  // namespace gvsk
