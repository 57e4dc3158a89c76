// Exact FP64 brute-force kNN (ground truth) on MFMA.
//
// The reference loads ground-truth files (INTER/host.cpp:148-192); hnswlib's
// BruteforceSearch (HNSWLIB/bruteforce.h) scans every point.  Here the scan
// is a dense FP64 GEMM: dist = |q|^2 + |x|^2 - 2 q.x (L2) or 1 - q.x (IP),
// q.x on v_mfma_f64_16x16x4f64, fused with a per-query top-k kept in LDS.
//
// Work split: a block of W waves owns 16*W queries and one chunk of the base
// set; each wave owns 16 queries.  Base tiles of 64 points are staged in LDS
// (fp32, odd row stride against bank conflicts) and widened to fp64 at the
// MFMA operand.  Each chunk leaves a sorted partial top-k per query; a second
// kernel merges the chunks.  Order: (dist, id) ascending.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "gvs_internal.hpp"

namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int WAVE = 64;
constexpr int TILE = 64;  // base points per LDS tile

__device__ __forceinline__ bool dlt(double a, int ia, double b, int ib) { return a < b || (a == b && ia < ib); }

__device__ __forceinline__ void lds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_xnorm64(const float *X, int64_t n, int dp, int dim, double *xn)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = X + i * dp;
    double s = 0.0;
    for (int d = 0; d < dim; d++) s += (double)r[d] * (double)r[d];
    xn[i] = s;
}

// LDS per block: Xs[TILE][dp+1] floats, then per wave lists ld[16][k] double, li[16][k] int
__global__ __launch_bounds__(256) void k_bf64(const float *X, const double *xn, int64_t n, int dp, int dim,
                                              const float *Q, int64_t ldq, int64_t nq, int k, int metric,
                                              int64_t chunk, double *part_d, int32_t *part_i)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int nw = blockDim.x / WAVE;
    const int w = threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int xs_stride = dp + 1;
    float *Xs = (float *)smem;
    double *ld = (double *)(smem + (((size_t)TILE * xs_stride * 4 + 15) & ~(size_t)15)) + (size_t)w * 16 * k;
    int32_t *li = (int32_t *)((double *)(smem + (((size_t)TILE * xs_stride * 4 + 15) & ~(size_t)15)) +
                              (size_t)nw * 16 * k) + (size_t)w * 16 * k;

    const int64_t q0 = ((int64_t)blockIdx.x * nw + w) * 16;
    const int64_t c0 = (int64_t)blockIdx.y * chunk;
    const int64_t c1 = min(n, c0 + chunk);

    // lists start empty
    for (int i = lane; i < 16 * k; i += WAVE) {
        ld[i] = __builtin_inf();
        li[i] = 0x7FFFFFFF;
    }
    // this lane's query operand row (A[i][kk] = Q[q0 + l%16][4 kk + l/16])
    const int ai = lane & 15, ak = lane >> 4;
    const int64_t qa = q0 + ai;
    const float *qrow = Q + min(qa, nq - 1) * ldq;
    lds_wave_sync();
    // Output layout of v_mfma_f64_16x16x4f64, read from the instruction
    // itself: with A[i][0] = i, B[0][j] = 1 every D[i][j] equals its row i;
    // with A[i][0] = 1, B[0][j] = j it equals its column j.
    int rowid[4], colid[4];
    {
        const double4_t z = (double4_t){0.0, 0.0, 0.0, 0.0};
        double4_t pr = __builtin_amdgcn_mfma_f64_16x16x4f64(ak == 0 ? (double)ai : 0.0, ak == 0 ? 1.0 : 0.0, z, 0, 0, 0);
        double4_t pc = __builtin_amdgcn_mfma_f64_16x16x4f64(ak == 0 ? 1.0 : 0.0, ak == 0 ? (double)ai : 0.0, z, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            rowid[r] = min(15, max(0, (int)pr[r]));
            colid[r] = min(15, max(0, (int)pc[r]));
        }
    }
    // query norms of the rows this lane holds in D
    double qn[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int64_t qi = q0 + rowid[r];
        const float *qq = Q + min(qi, nq - 1) * ldq;
        double s = 0.0;
        for (int d = 0; d < dim; d++) s += (double)qq[d] * (double)qq[d];
        qn[r] = s;
    }
    __syncthreads();

    for (int64_t t0 = c0; t0 < c1; t0 += TILE) {
        // stage TILE points (fp32) into LDS
        for (int e = threadIdx.x; e < TILE * dp; e += blockDim.x) {
            int pnt = e / dp, col = e - pnt * dp;
            int64_t id = t0 + pnt;
            Xs[pnt * xs_stride + col] = (id < c1) ? X[id * dp + col] : 0.0f;
        }
        __syncthreads();
        if (q0 < nq) {
            double4_t acc[4];
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = (double4_t){0.0, 0.0, 0.0, 0.0};
            for (int kk = 0; kk < dp / 4; kk++) {
                int col = 4 * kk + ak;
                double a = (col < dim) ? (double)qrow[col] : 0.0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    double b = (double)Xs[(16 * c + ai) * xs_stride + col];
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
                }
            }
            // D[i][j] of tile c: this lane holds rows rowid[r], column colid[r];
            // for a fixed r the 16 lanes of a quarter-wave share one row
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = 16 * c + colid[r];
                    const int64_t id64 = t0 + j;
                    const bool valid = id64 < c1;
                    const int id = (int)id64;
                    const double xnj = valid ? xn[id64] : 0.0;
                    const double dot = acc[c][r];
                    const double d = (metric == 0) ? (qn[r] + xnj - 2.0 * dot) : (1.0 - dot);
                    const int own_row = rowid[r];
                    // survivor: beats the current k-th entry (list tail, +inf while not full)
                    bool s = valid && dlt(d, id, ld[own_row * k + k - 1], li[own_row * k + k - 1]);
                    uint64_t bal = __ballot(s);
                    while (bal) {
                        // one survivor per 16-lane group (= per query row) per round
                        uint64_t pick = 0;
#pragma unroll
                        for (int g = 0; g < 4; g++) {
                            uint64_t m = (bal >> (16 * g)) & 0xFFFFull;
                            if (m) pick |= (m & (~m + 1)) << (16 * g);
                        }
#pragma unroll
                        for (int g = 0; g < 4; g++) {
                            const uint64_t m = (pick >> (16 * g)) & 0xFFFFull;
                            if (!m) continue;
                            const int src = __ffsll((unsigned long long)m) - 1 + 16 * g;
                            const double nd = __shfl(d, src);
                            const int nid = __shfl(id, src);
                            const int row = __shfl(own_row, src);
                            double *L = ld + row * k;
                            int32_t *I = li + row * k;
                            int pos = 0;
                            for (int b = 0; b < k; b += WAVE) {
                                int j2 = b + lane;
                                bool lt = j2 < k && dlt(L[j2], I[j2], nd, nid);
                                pos += __popcll(__ballot(lt));
                            }
                            if (pos < k) {
                                // shift [pos, k-1) up by one, top slab first
                                for (int b = ((k - 1) / WAVE) * WAVE; b >= 0; b -= WAVE) {
                                    int j2 = b + lane;
                                    bool mv = j2 >= pos && j2 + 1 < k;
                                    double vd = mv ? L[j2] : 0.0;
                                    int vi = mv ? I[j2] : 0;
                                    lds_wave_sync();
                                    if (mv) {
                                        L[j2 + 1] = vd;
                                        I[j2 + 1] = vi;
                                    }
                                    lds_wave_sync();
                                }
                                if (lane == 0) {
                                    L[pos] = nd;
                                    I[pos] = nid;
                                }
                                lds_wave_sync();
                            }
                        }
                        s = s && !((pick >> lane) & 1ull) &&
                            dlt(d, id, ld[own_row * k + k - 1], li[own_row * k + k - 1]);
                        bal = __ballot(s);
                    }
                }
            }
        }
        __syncthreads();
    }
    // write partial lists
    if (q0 < nq) {
        for (int row = 0; row < 16; row++) {
            int64_t qi = q0 + row;
            if (qi >= nq) break;
            for (int j2 = lane; j2 < k; j2 += WAVE) {
                size_t o = ((size_t)blockIdx.y * nq + qi) * k + j2;
                part_d[o] = ld[row * k + j2];
                part_i[o] = li[row * k + j2] == 0x7FFFFFFF ? -1 : li[row * k + j2];
            }
        }
    }
}

// v2 for padded dims of at most 128 (NK = dp / 4 k-steps): the wave's 16
// query rows are held in registers as fp64 operands for the whole chunk
// (A[i][kk] = Q[q0 + l%16][4 kk + l/16]); base tiles are double-buffered in
// LDS (fp32) with their norms, the next tile's global loads in flight while
// the current one is multiplied; one barrier per tile.
//
// Tile layout: the MFMA of k-step kk and column block c reads B[4 kk + ak]
// [16 c + ai] on lane ai + 16 ak.  Element (point 16 c + ai, column 4 kk +
// ak) sits at kk * KSTEP + 64 c + 32 (ak >> 1) + 2 ai + (ak & 1), so each
// half-wave reads 32 consecutive floats (no bank conflicts); KSTEP = 258
// spreads the staging writes of one point's k-steps over the banks.
constexpr int KSTEP = 4 * TILE + 2;
#ifndef GVS_BF_PREFETCH
#define GVS_BF_PREFETCH 2  // k-steps of B operands loaded ahead of their products
#endif
//
// Top-k lists: with REGK (k <= 16) each row's list lives in registers of the
// 16 lanes that hold the row in D (lane ai keeps entry ai), and a survivor
// is inserted by shuffles within the group, the four groups in parallel;
// otherwise each row's list is in LDS and survivors go in one at a time.
template <int NK, bool REGK>
__global__ __launch_bounds__(256, 2) void k_bf64_v2(const float *X, const double *xn, int64_t n, int dim,
                                                 const float *Q, int64_t ldq, int64_t nq, int k, int metric,
                                                 int64_t chunk, double *part_d, int32_t *part_i)
{
    constexpr int DP = 4 * NK;
    constexpr int XS = KSTEP * NK / TILE + 1;  // tile floats = TILE * XS >= NK * KSTEP
    static_assert(TILE * XS >= NK * KSTEP, "tile buffer too small");
    constexpr int F4 = DP / 4;                 // float4 per base row
    constexpr int PER = TILE * F4 / 256;       // float4 per thread per tile
    static_assert(TILE * F4 % 256 == 0, "tile must split evenly over 256 threads");
    extern __shared__ __align__(16) unsigned char smem[];
    const int nw = blockDim.x / WAVE;
    const int w = threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int tid = threadIdx.x;
    float *const Xs0 = (float *)smem;  // tile b at Xs0 + b * TILE * XS
    double *const xs_n0 = (double *)(smem + (((size_t)2 * TILE * XS * 4 + 15) & ~(size_t)15));  // norms at + b * TILE
    double *ld = xs_n0 + 2 * TILE + (size_t)w * 16 * k;
    int32_t *li = (int32_t *)(xs_n0 + 2 * TILE + (size_t)nw * 16 * k) + (size_t)w * 16 * k;

    const int64_t q0 = ((int64_t)blockIdx.x * nw + w) * 16;
    const int64_t c0 = (int64_t)blockIdx.y * chunk;
    const int64_t c1 = min(n, c0 + chunk);
    if (!REGK) {
        for (int i = lane; i < 16 * k; i += WAVE) {
            ld[i] = __builtin_inf();
            li[i] = 0x7FFFFFFF;
        }
    }
    const int ai = lane & 15, ak = lane >> 4;
    // query operands, fp64, for the whole chunk
    // (kept as fp32: widening is exact, and it halves the registers)
    float areg[NK];
    {
        const int64_t qa = q0 + ai;
        const float *qrow = Q + min(qa, nq - 1) * ldq;
#pragma unroll
        for (int kk = 0; kk < NK; kk++) {
            const int col = 4 * kk + ak;
            areg[kk] = (col < dim && qa < nq) ? qrow[col] : 0.0f;
        }
    }
    // D of v_mfma_f64_16x16x4f64 on gfx950: element r of lane l is
    // D[l / 16 + 4 r][l % 16] (the probe in k_bf64 reads the same from the
    // instruction), so a 16-lane group holds one row per r, all 16 columns
    int rowid[4], colid[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        rowid[r] = ak + 4 * r;
        colid[r] = ai;
    }
    // REGK lists: entry ai of row ak + 4 r in rl_d[r] / rl_i[r]; thr_* is
    // entry k - 1 (the bound a candidate must beat), kept in every lane
    double rl_d[4], thr_d[4];
    int rl_i[4], thr_i[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        rl_d[r] = thr_d[r] = __builtin_inf();
        rl_i[r] = thr_i[r] = 0x7FFFFFFF;
    }
    double qn[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float *qq = Q + min(q0 + rowid[r], nq - 1) * ldq;
        double sq = 0.0;
        for (int d = 0; d < dim; d++) sq += (double)qq[d] * (double)qq[d];
        qn[r] = sq;
    }
    const float4 *X4 = (const float4 *)X;
    float4 stg[PER];
    double stg_n = 0.0;
    auto load_tile = [&](int64_t t0) {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int e = tid + 256 * j, pnt = e / F4, c4 = e - pnt * F4;
            const int64_t id = t0 + pnt;
            stg[j] = id < c1 ? X4[id * F4 + c4] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (tid < TILE) stg_n = t0 + tid < c1 ? xn[t0 + tid] : 0.0;
    };
    auto store_tile = [&](int b) {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int e = tid + 256 * j, pnt = e / F4, c4 = e - pnt * F4;
            float *dst = Xs0 + (size_t)b * TILE * XS + c4 * KSTEP + 64 * (pnt >> 4) + 2 * (pnt & 15);
            *(float2 *)dst = make_float2(stg[j].x, stg[j].y);          // columns 4 c4 + 0, 1
            *(float2 *)(dst + 32) = make_float2(stg[j].z, stg[j].w);   // columns 4 c4 + 2, 3
        }
        if (tid < TILE) xs_n0[b * TILE + tid] = stg_n;
    };
    load_tile(c0);
    store_tile(0);
    __syncthreads();
    int buf = 0;
    for (int64_t t0 = c0; t0 < c1; t0 += TILE, buf ^= 1) {
        const bool more = t0 + TILE < c1;
        if (more) load_tile(t0 + TILE);  // in flight while this tile is multiplied
        if (q0 < nq) {
            const float *B = Xs0 + (size_t)buf * TILE * XS;
            double4_t acc[4];
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = (double4_t){0.0, 0.0, 0.0, 0.0};
            const float *Bl = B + 32 * (ak >> 1) + 2 * ai + (ak & 1);
            // B values run PF k-steps ahead of the products that use them
            constexpr int PF = GVS_BF_PREFETCH;
            float bq[PF + 1][4];
#pragma unroll
            for (int p = 0; p < PF; p++)
#pragma unroll
                for (int c = 0; c < 4; c++) bq[p][c] = Bl[p * KSTEP + 64 * c];
#pragma unroll
            for (int kk = 0; kk < NK; kk++) {
                if (kk + PF < NK) {
#pragma unroll
                    for (int c = 0; c < 4; c++) bq[(kk + PF) % (PF + 1)][c] = Bl[(kk + PF) * KSTEP + 64 * c];
                }
                // keep the loads ahead: the scheduler would otherwise sink
                // each one next to its product and wait on it there
                __builtin_amdgcn_sched_barrier(0);
                const double a = (double)areg[kk];
#pragma unroll
                for (int c = 0; c < 4; c++)
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, (double)bq[kk % (PF + 1)][c], acc[c], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (REGK) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int j = 16 * c + ai;
                    const int64_t id64 = t0 + j;
                    const bool valid = id64 < c1;
                    const int id = (int)id64;
                    const double xnj = xs_n0[buf * TILE + j];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const double dot = acc[c][r];
                        const double d = (metric == 0) ? (qn[r] + xnj - 2.0 * dot) : (1.0 - dot);
                        bool sv = valid && dlt(d, id, thr_d[r], thr_i[r]);
                        uint64_t bal = __ballot(sv);
                        while (bal) {
                            // each group takes its lowest survivor
                            const uint32_t m = (uint32_t)(bal >> (16 * ak)) & 0xFFFFu;
                            const int src = m ? 16 * ak + __ffs((int)m) - 1 : lane;
                            const double nd = __shfl(d, src);
                            const int nid = __shfl(id, src);
                            const bool lt = ai < k && dlt(rl_d[r], rl_i[r], nd, nid);
                            const int pos = __popcll((__ballot(lt) >> (16 * ak)) & 0xFFFFull);
                            const double pd = __shfl(rl_d[r], max(lane - 1, 0));
                            const int pi = __shfl(rl_i[r], max(lane - 1, 0));
                            if (m && ai >= pos) {
                                rl_d[r] = ai == pos ? nd : pd;
                                rl_i[r] = ai == pos ? nid : pi;
                            }
                            thr_d[r] = __shfl(rl_d[r], 16 * ak + k - 1);
                            thr_i[r] = __shfl(rl_i[r], 16 * ak + k - 1);
                            sv = sv && lane != src && dlt(d, id, thr_d[r], thr_i[r]);
                            bal = __ballot(sv);
                        }
                    }
                }
            } else {
            // the rows' current k-th entries (a list only tightens, so a
            // stale bound lets a few more candidates into the exact check)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                thr_d[r] = ld[rowid[r] * k + k - 1];
                thr_i[r] = li[rowid[r] * k + k - 1];
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = 16 * c + colid[r];
                    const int64_t id64 = t0 + j;
                    const bool valid = id64 < c1;
                    const int id = (int)id64;
                    const double dot = acc[c][r];
                    const double d = (metric == 0) ? (qn[r] + xs_n0[buf * TILE + j] - 2.0 * dot) : (1.0 - dot);
                    const int own_row = rowid[r];
                    bool sv = valid && dlt(d, id, thr_d[r], thr_i[r]);
                    uint64_t bal = __ballot(sv);
                    while (bal) {
                        uint64_t pick = 0;
#pragma unroll
                        for (int g = 0; g < 4; g++) {
                            uint64_t m = (bal >> (16 * g)) & 0xFFFFull;
                            if (m) pick |= (m & (~m + 1)) << (16 * g);
                        }
#pragma unroll
                        for (int g = 0; g < 4; g++) {
                            const uint64_t m = (pick >> (16 * g)) & 0xFFFFull;
                            if (!m) continue;
                            const int src = __ffsll((unsigned long long)m) - 1 + 16 * g;
                            const double nd = __shfl(d, src);
                            const int nid = __shfl(id, src);
                            const int row = __shfl(own_row, src);
                            double *L = ld + row * k;
                            int32_t *I = li + row * k;
                            int pos = 0;
                            for (int b = 0; b < k; b += WAVE) {
                                int j2 = b + lane;
                                bool lt = j2 < k && dlt(L[j2], I[j2], nd, nid);
                                pos += __popcll(__ballot(lt));
                            }
                            if (pos < k) {
                                for (int b = ((k - 1) / WAVE) * WAVE; b >= 0; b -= WAVE) {
                                    int j2 = b + lane;
                                    bool mv = j2 >= pos && j2 + 1 < k;
                                    double vd = mv ? L[j2] : 0.0;
                                    int vi = mv ? I[j2] : 0;
                                    lds_wave_sync();
                                    if (mv) {
                                        L[j2 + 1] = vd;
                                        I[j2 + 1] = vi;
                                    }
                                    lds_wave_sync();
                                }
                                if (lane == 0) {
                                    L[pos] = nd;
                                    I[pos] = nid;
                                }
                                lds_wave_sync();
                            }
                        }
                        sv = sv && !((pick >> lane) & 1ull) &&
                             dlt(d, id, ld[own_row * k + k - 1], li[own_row * k + k - 1]);
                        bal = __ballot(sv);
                    }
                }
            }
            }
        }
        if (more) store_tile(buf ^ 1);  // nobody reads buf ^ 1 until after the barrier
        __syncthreads();
    }
    if (REGK) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t qi = q0 + rowid[r];
            if (qi < nq && ai < k) {
                const size_t o = ((size_t)blockIdx.y * nq + qi) * k + ai;
                part_d[o] = rl_d[r];
                part_i[o] = rl_i[r] == 0x7FFFFFFF ? -1 : rl_i[r];
            }
        }
    } else if (q0 < nq) {
        for (int row = 0; row < 16; row++) {
            int64_t qi = q0 + row;
            if (qi >= nq) break;
            for (int j2 = lane; j2 < k; j2 += WAVE) {
                size_t o = ((size_t)blockIdx.y * nq + qi) * k + j2;
                part_d[o] = ld[row * k + j2];
                part_i[o] = li[row * k + j2] == 0x7FFFFFFF ? -1 : li[row * k + j2];
            }
        }
    }
}

typedef void (*bf64v2_t)(const float *, const double *, int64_t, int, const float *, int64_t, int64_t, int, int,
                         int64_t, double *, int32_t *);
constexpr int REGK_MAX = 16;  // lists in registers up to this k
static bf64v2_t pick_v2(int dp, int k)
{
    const bool rk = k <= REGK_MAX;
    switch (dp) {
    case 32: return rk ? k_bf64_v2<8, true> : k_bf64_v2<8, false>;
    case 64: return rk ? k_bf64_v2<16, true> : k_bf64_v2<16, false>;
    case 96: return rk ? k_bf64_v2<24, true> : k_bf64_v2<24, false>;
    case 128: return rk ? k_bf64_v2<32, true> : k_bf64_v2<32, false>;
    default: return nullptr;
    }
}

// merge nchunks sorted partial lists per query (one wave per query)
__global__ void k_bf64_merge(const double *part_d, const int32_t *part_i, int nchunks, int64_t nq, int k,
                             int32_t *out_i, double *out_d)
{
    const int64_t qi = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (qi >= nq) return;
    const int total = nchunks * k;
    for (int e = lane; e < total; e += WAVE) {
        int c = e / k, j = e - c * k;
        size_t o = ((size_t)c * nq + qi) * k + j;
        double d = part_d[o];
        int id = part_i[o];
        if (id < 0) continue;
        // rank among all candidates; each list is sorted so a binary search per
        // list gives the count of smaller entries
        int rank = 0;
        for (int c2 = 0; c2 < nchunks; c2++) {
            const double *Ld = part_d + ((size_t)c2 * nq + qi) * k;
            const int32_t *Li = part_i + ((size_t)c2 * nq + qi) * k;
            int lo = 0, hi = k;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                bool lt = Li[mid] >= 0 && dlt(Ld[mid], Li[mid], d, id);
                if (lt) lo = mid + 1;
                else hi = mid;
            }
            rank += lo;
        }
        if (rank < k) {
            out_i[qi * k + rank] = id;
            out_d[qi * k + rank] = d;
        }
    }
}

__global__ void k_fill_missing(int32_t *out_i, double *out_d, int64_t total)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) {
        out_i[i] = -1;
        out_d[i] = __builtin_inf();
    }
}

}  // namespace

extern "C" int gvs_bruteforce_f64_device(const gvs_index *cix, const float *queries, int64_t nq, int64_t ldq,
                                         int32_t k, int32_t *ids, double *dists, void *stream_)
{
    gvs_index *ix = const_cast<gvs_index *>(cix);
    hipStream_t stream = (hipStream_t)stream_;
    if (!ix || nq < 0 || k < 1 || k > 1024 || ldq < ix->dim || (nq > 0 && (!queries || !ids || !dists))) {
        gvs::set_error("gvs_bruteforce_f64_device: invalid argument");
        return GVS_ERR_ARG;
    }
    if (nq == 0) return GVS_OK;
    const int dp = ix->dim_pad;
    if (!ix->d_xnorm64) {
        if (!gvs::hip_ok(hipMalloc(&ix->d_xnorm64, sizeof(double) * ix->n), "hipMalloc xnorm")) return GVS_ERR_NOMEM;
        hipLaunchKernelGGL(k_xnorm64, dim3((unsigned)((ix->n + 255) / 256)), dim3(256), 0, stream, ix->d_f32, ix->n,
                           dp, ix->dim, ix->d_xnorm64);
    }
    const size_t per_wave = (size_t)16 * k * 12;
    // v2 (register-resident queries, double-buffered tiles) where the padded
    // dim is compiled in and four waves' lists fit beside two tiles
    bf64v2_t v2 = pick_v2(dp, k);
    const int xs_v2 = KSTEP * (dp / 4) / TILE + 1;  // as the kernel's XS
    const size_t xs2 = (((size_t)2 * TILE * xs_v2 * 4 + 15) & ~(size_t)15) + (size_t)2 * TILE * 8;
    const size_t v2_lists = k <= REGK_MAX ? 0 : 4 * per_wave;  // register lists need no LDS
    if (v2 && xs2 + v2_lists > 160 * 1024) v2 = nullptr;
    const size_t xs_bytes = v2 ? xs2 : (((size_t)TILE * (dp + 1) * 4 + 15) & ~(size_t)15);
    int wpb = 4;
    while (!v2 && wpb > 1 && xs_bytes + wpb * per_wave > 160 * 1024) wpb--;
    const size_t shm = v2 ? xs2 + v2_lists : xs_bytes + wpb * per_wave;
    if (shm > 160 * 1024) {
        gvs::set_error("gvs_bruteforce_f64_device: k=%d too large for LDS at dim %d", k, ix->dim);
        return GVS_ERR_UNSUPPORTED;
    }
    const void *kfn = v2 ? (const void *)v2 : (const void *)k_bf64;
    if (shm > 65536 && !gvs::hip_ok(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                                    "hipFuncSetAttribute"))
        return GVS_ERR_HIP;
    int bpc = 1;
    if (!gvs::hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kfn, wpb * WAVE, shm), "occupancy"))
        return GVS_ERR_HIP;
    const int64_t resident = (int64_t)std::max(1, bpc) * ix->cu_count;
    const int64_t qblocks = (nq + 16 * wpb - 1) / (16 * wpb);
    // chunks of the base set: enough blocks to fill the GPU, the count
    // chosen so the last round of resident blocks is as full as possible
    const int64_t max_chunks = std::max<int64_t>(1, (ix->n + TILE * 16 - 1) / (TILE * 16));
    int64_t nchunks = 1;
    double best_eff = -1.0;
    const int64_t c_lo = std::max<int64_t>(1, std::min<int64_t>(max_chunks, (resident + qblocks - 1) / qblocks / 2));
    const int64_t c_hi = std::max<int64_t>(c_lo, std::min<int64_t>(max_chunks, 4 * resident / qblocks + 1));
    for (int64_t c = c_lo; c <= c_hi; c++) {
        const int64_t blocks = qblocks * c;
        const int64_t rounds = (blocks + resident - 1) / resident;
        const double eff = (double)blocks / (double)(rounds * resident);
        if (eff > best_eff + 1e-9) {
            best_eff = eff;
            nchunks = c;
        }
    }
    int64_t chunk = (ix->n + nchunks - 1) / nchunks;
    chunk = (chunk + TILE - 1) / TILE * TILE;
    nchunks = (ix->n + chunk - 1) / chunk;
    double *pd = nullptr;
    int32_t *pi = nullptr;
    if (!gvs::hip_ok(hipMallocAsync((void **)&pd, sizeof(double) * nchunks * nq * k, stream), "hipMallocAsync") ||
        !gvs::hip_ok(hipMallocAsync((void **)&pi, sizeof(int32_t) * nchunks * nq * k, stream), "hipMallocAsync"))
        return GVS_ERR_NOMEM;
    if (v2)
        hipLaunchKernelGGL(v2, dim3((unsigned)qblocks, (unsigned)nchunks), dim3(wpb * WAVE), shm, stream, ix->d_f32,
                           ix->d_xnorm64, ix->n, ix->dim, queries, ldq, nq, k, ix->metric, chunk, pd, pi);
    else
        hipLaunchKernelGGL(k_bf64, dim3((unsigned)qblocks, (unsigned)nchunks), dim3(wpb * WAVE), shm, stream,
                           ix->d_f32, ix->d_xnorm64, ix->n, dp, ix->dim, queries, ldq, nq, k, ix->metric, chunk, pd, pi);
    if (!gvs::hip_ok(hipGetLastError(), "launch k_bf64")) return GVS_ERR_HIP;
    hipLaunchKernelGGL(k_fill_missing, dim3((unsigned)((nq * k + 255) / 256)), dim3(256), 0, stream, ids, dists,
                       nq * k);
    hipLaunchKernelGGL(k_bf64_merge, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, stream, pd, pi, (int)nchunks, nq, k,
                       ids, dists);
    if (!gvs::hip_ok(hipGetLastError(), "launch k_bf64_merge")) return GVS_ERR_HIP;
    (void)hipFreeAsync(pd, stream);
    (void)hipFreeAsync(pi, stream);
    return GVS_OK;
}
