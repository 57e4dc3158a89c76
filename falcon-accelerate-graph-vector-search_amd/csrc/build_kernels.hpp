// Layer-0 construction kernels of the GPU builder (hnswlib's insertion,
// networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h, run for a batch of
// new nodes at once against the graph as it stood at the batch start).
//
//   k_build_select  one wave per new node: getNeighborsByHeuristic2
//                   (:328-369) over the node's construction-search
//                   candidates, keeping M; writes the node's row farthest
//                   first (mutuallyConnectNewElement, :386-440) and one
//                   reverse-link key (target << 32 | node) per selected
//                   neighbour.
//   k_build_link    one wave per run of sorted reverse-link keys with one
//                   target: the target's row takes the new links in node
//                   order when they fit (:466-470), otherwise it is re-pruned
//                   with the heuristic over its links and the new ones
//                   (:471-497, all new links at once).
//
// Distances are the search kernels' (the fp32 8-lane order, or hnswlib's
// scalar order in the STORE_F32_SEQ instantiation), so a batch of one node
// with the scalar order reproduces hnswlib's graph.
#pragma once
#include "search_query.hpp"

namespace gvsk {

constexpr int BUILD_CAP = 1024;  // candidates per heuristic call (efConstruction, or row + new links)

struct BuildArgs {
    int64_t first;          // first node of the batch
    int nb;                 // nodes in the batch
    int efc;                // candidates per node (row width of cand / cdist)
    const int64_t *cand;    // nb x efc search results, ascending, -1 padded
    const float *cdist;     // their distances
    int M;                  // neighbours a new node keeps
    int32_t *adj0;          // n x r0 rows (P.r0), compact, -1 padded
    uint64_t *pairs;        // nb x M reverse-link keys (~0 = none), sorted between the two kernels
    int64_t npairs;         // k_build_link: keys in pairs
};

struct BuildLds {
    uint64_t keys[BUILD_CAP];  // candidate keys (order), then scratch
    float dist[BUILD_CAP];
    uint32_t ids[BUILD_CAP];
    uint64_t tk[WAVE];         // distances from one candidate to the kept ones
    uint32_t sel[WAVE];
    float seld[WAVE];
};

// The candidate's vector as a search query (lane p = lane & 7 holds the
// elements (p + 8 i) * E + t, the layout eval_rows reads).
template <int NCH>
__device__ __forceinline__ void load_row_as_query(const KParams &P, uint32_t id, float *qf)
{
    const int p = lane_id() & 7;
    const float *row = (const float *)P.vec_f32 + (size_t)id * (size_t)P.row_elems_f32;
#pragma unroll
    for (int i = 0; i < NCH; i++) {
        const float4 v = *(const float4 *)(row + (p + 8 * i) * 4);
        qf[i * 4 + 0] = v.x;
        qf[i * 4 + 1] = v.y;
        qf[i * 4 + 2] = v.z;
        qf[i * 4 + 3] = v.w;
    }
}

// getNeighborsByHeuristic2 (:328-369) over S.ids / S.dist [0, nc), which
// are in the order hnswlib pops them (distance ascending, equal distances
// larger id first): a candidate is kept unless a kept one is closer to it
// than it is to the base point; at most m are kept; fewer than m candidates
// are all kept.  Kept ids and distances go to S.sel / S.seld, in keeping
// order.  Returns how many.
template <int NCH, int STORE, int METRIC>
__device__ int heuristic_keep(const KParams &P, BuildLds &S, int nc, int m)
{
    const int lane = lane_id();
    nc = uni(nc);
    if (nc < m) {
        for (int i = lane; i < nc; i += WAVE) {
            S.sel[i] = S.ids[i];
            S.seld[i] = S.dist[i];
        }
        wave_sync();
        return nc;
    }
    int nsel = 0;
    float qf[NCH * 4];
    for (int c = 0; c < nc && nsel < m; c++) {
        const uint32_t id = S.ids[c];
        const float dq = S.dist[c];
        bool good = true;
        if (nsel > 0) {
            load_row_as_query<NCH>(P, id, qf);
            eval_rows<NCH, STORE, METRIC>(P, qf, S.sel, nsel, S.tk);
            wave_sync();
            const bool closer = lane < nsel && key_dist(S.tk[lane]) < dq;
            good = __ballot(closer) == 0ull;
        }
        if (good) {
            if (lane == 0) {
                S.sel[nsel] = id;
                S.seld[nsel] = dq;
            }
            nsel++;
            wave_sync();
        }
    }
    return nsel;
}

// Row of `node` from the kept neighbours: farthest first (the order hnswlib's
// max-heap pops (distance, id) pairs in), -1 padded to r0.
__device__ __forceinline__ void write_kept_row(const BuildLds &S, int nsel, int32_t *row, int r0)
{
    const int lane = lane_id();
    if (lane < nsel) {
        const float d = S.seld[lane];
        const uint32_t id = S.sel[lane];
        int rank = 0;
        for (int j = 0; j < nsel; j++) {
            const float dj = S.seld[j];
            const uint32_t ij = S.sel[j];
            rank += (dj > d || (dj == d && ij > id)) ? 1 : 0;
        }
        row[rank] = (int32_t)id;
    }
    for (int j = nsel + lane; j < r0; j += WAVE) row[j] = -1;
}

// Reorders S.ids / S.dist [0, nc), sorted by (distance, id), so that equal
// distances come larger id first (hnswlib's pop order).
__device__ __forceinline__ void ties_larger_id_first(BuildLds &S, int nc)
{
    const int lane = lane_id();
    bool tie = false;
    for (int i = lane; i + 1 < nc; i += WAVE) tie |= S.dist[i] == S.dist[i + 1];
    if (!__ballot(tie)) return;
    for (int i = lane; i < nc; i += WAVE) {
        const float d = S.dist[i];
        int s = i, e = i + 1;
        while (s > 0 && S.dist[s - 1] == d) s--;
        while (e < nc && S.dist[e] == d) e++;
        S.keys[s + e - 1 - i] = ((uint64_t)__float_as_uint(d) << 32) | S.ids[i];
    }
    wave_sync();
    for (int i = lane; i < nc; i += WAVE) {
        S.ids[i] = (uint32_t)S.keys[i];
        S.dist[i] = __uint_as_float((uint32_t)(S.keys[i] >> 32));
    }
    wave_sync();
}

template <int NCH, int STORE, int METRIC>
__global__ __launch_bounds__(64) void k_build_select(KParams P, BuildArgs A)
{
    __shared__ BuildLds S;
    const int b = blockIdx.x;
    if (b >= A.nb) return;
    const int lane = lane_id();
    const uint32_t node = (uint32_t)(A.first + b);
    const int64_t *cand = A.cand + (size_t)b * A.efc;
    const float *cd = A.cdist + (size_t)b * A.efc;
    int nc = 0;
    for (int i0 = 0; i0 < A.efc; i0 += WAVE) {
        const int i = i0 + lane;
        const bool v = i < A.efc && cand[i] >= 0;
        if (v) {
            S.ids[i] = (uint32_t)cand[i];
            S.dist[i] = cd[i];
        }
        nc += __popcll(__ballot(v));
    }
    wave_sync();
    ties_larger_id_first(S, nc);
    const int nsel = heuristic_keep<NCH, STORE, METRIC>(P, S, nc, A.M);
    write_kept_row(S, nsel, A.adj0 + (size_t)node * P.r0, P.r0);
    for (int j = lane; j < A.M; j += WAVE)
        A.pairs[(size_t)b * A.M + j] = j < nsel ? ((uint64_t)S.sel[j] << 32) | node : ~0ull;
}

// In-LDS bitonic sort of keys [0, n) ascending (n padded to a power of two
// with ~0 keys, n <= BUILD_CAP).
__device__ __forceinline__ void lds_sort(uint64_t *keys, int n)
{
    const int lane = lane_id();
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int i = n + lane; i < np2; i += WAVE) keys[i] = ~0ull;
    wave_sync();
    for (int k = 2; k <= np2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < np2; i += WAVE) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = keys[i], c = keys[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        keys[i] = c;
                        keys[ixj] = a;
                    }
                }
            }
            wave_sync();
        }
}

template <int NCH, int STORE, int METRIC>
__global__ __launch_bounds__(64) void k_build_link(KParams P, BuildArgs A)
{
    __shared__ BuildLds S;
    const int64_t i = blockIdx.x;
    if (i >= A.npairs) return;
    const uint64_t key = A.pairs[i];
    if (key == ~0ull) return;
    const uint32_t s = (uint32_t)(key >> 32);
    if (i > 0 && (uint32_t)(A.pairs[i - 1] >> 32) == s) return;  // not the first key of its target
    const int lane = lane_id();
    // the run of keys with target s
    int64_t j = i + 1;
    for (;;) {
        const int64_t t = j + lane;
        const bool same = t < A.npairs && A.pairs[t] != ~0ull && (uint32_t)(A.pairs[t] >> 32) == s;
        const uint64_t bal = __ballot(same);
        const int run = bal == ~0ull ? WAVE : (int)__builtin_ctzll(~bal);  // leading lanes with the same target
        j += run;
        if (run < WAVE) break;
    }
    const int r0 = P.r0;
    int32_t *row = A.adj0 + (size_t)s * r0;
    int c0 = 0;
    for (int b0 = 0; b0 < r0; b0 += WAVE) {
        const bool v = b0 + lane < r0 && row[b0 + lane] >= 0;
        c0 += __popcll(__ballot(v));
    }
    c0 = uni(c0);
    float qf[NCH * 4];
    bool have_q = false;
    for (int64_t t0 = i; t0 < j;) {
        const int m = (int)min<int64_t>(j - t0, (int64_t)(BUILD_CAP - r0));
        if (c0 + m <= r0) {  // room: append in node order
            for (int t = lane; t < m; t += WAVE) row[c0 + t] = (int32_t)(uint32_t)A.pairs[t0 + t];
            c0 += m;
            wave_sync();
        } else {
            // re-prune: the row's links and the new ones, by distance to s
            const int nc = c0 + m;
            for (int t = lane; t < nc; t += WAVE)
                S.ids[t] = t < c0 ? (uint32_t)row[t] : (uint32_t)A.pairs[t0 + (t - c0)];
            wave_sync();
            if (!have_q) {
                load_row_as_query<NCH>(P, s, qf);
                have_q = true;
            }
            eval_rows<NCH, STORE, METRIC>(P, qf, S.ids, nc, S.keys);
            wave_sync();
            // order: distance ascending, equal distances larger id first
            for (int t = lane; t < nc; t += WAVE) {
                const uint64_t k = S.keys[t];
                S.keys[t] = (k & 0xFFFFFFFF00000000ull) | (uint64_t)(0xFFFFFFFFu - key_id(k));
            }
            wave_sync();
            lds_sort(S.keys, nc);
            for (int t = lane; t < nc; t += WAVE) {
                const uint64_t k = S.keys[t];
                S.ids[t] = 0xFFFFFFFFu - (uint32_t)k;
                S.dist[t] = unord_f((uint32_t)(k >> 32));
            }
            wave_sync();
            const int nsel = heuristic_keep<NCH, STORE, METRIC>(P, S, nc, r0);
            write_kept_row(S, nsel, row, r0);
            wave_sync();
            c0 = nsel;
        }
        t0 += m;
    }
}

}  // namespace gvsk
