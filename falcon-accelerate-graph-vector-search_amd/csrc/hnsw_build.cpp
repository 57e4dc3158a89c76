// HNSW graph construction with hnswlib's insertion semantics, multithreaded.
//
// Follows networked_FPGA/CPU_programs_oldest/hnswlib/hnswalg.h (HNSWLIB):
//   constructor fields (M, maxM = M, maxM0 = 2M, efc = max(efc, M),
//     mult = 1/ln(M), level RNG seeded with random_seed)   :27-66
//   getRandomLevel  -ln(U(0,1)) * mult                      :148-152
//   searchBaseLayer (construction search, efc)              :153-236
//   getNeighborsByHeuristic2                                :328-369
//   mutuallyConnectNewElement                               :386-501
//   addPoint (descent, per-level search + connect)          :982-1112
// Levels are drawn in index order before the parallel insertion, so one
// thread reproduces hnswlib's single-thread graph draw for draw.  Locks guard
// each node's lists; unlike hnswlib no lock is held across another lock.
// Priority queues are std::priority_queue with the same comparators, so the
// order of equal-distance pops matches hnswlib's as well.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <queue>
#include <random>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "host_graph.hpp"

namespace {

typedef std::pair<float, uint32_t> pf;
struct CmpFirst {
    bool operator()(const pf &a, const pf &b) const noexcept { return a.first < b.first; }
};
typedef std::priority_queue<pf, std::vector<pf>, CmpFirst> pq_t;

// hnswlib scalar order (space_l2.h:6-20 / space_ip.h:7-19); built with
// -ffp-contract=off so the multiply and add round separately.
float dist_exact(const float *a, const float *b, int d, int metric)
{
    float res = 0.0f;
    if (metric == GVS_METRIC_L2) {
        for (int i = 0; i < d; i++) {
            float t = a[i] - b[i];
            res += t * t;
        }
        return res;
    }
    for (int i = 0; i < d; i++) res += a[i] * b[i];
    return 1.0f - res;
}

// 16 independent partial sums (vectorises to AVX2 without reassociation)
float dist_fast(const float *a, const float *b, int d, int metric)
{
    float s[16] = {0};
    int i = 0;
    if (metric == GVS_METRIC_L2) {
        for (; i + 16 <= d; i += 16)
            for (int j = 0; j < 16; j++) {
                float t = a[i + j] - b[i + j];
                s[j] += t * t;
            }
        float r = 0.0f;
        for (; i < d; i++) {
            float t = a[i] - b[i];
            r += t * t;
        }
        for (int j = 0; j < 16; j++) r += s[j];
        return r;
    }
    for (; i + 16 <= d; i += 16)
        for (int j = 0; j < 16; j++) s[j] += a[i + j] * b[i + j];
    float r = 0.0f;
    for (; i < d; i++) r += a[i] * b[i];
    for (int j = 0; j < 16; j++) r += s[j];
    return 1.0f - r;
}

struct Builder {
    const float *x;
    int64_t n;
    int dim, metric;
    bool exact;
    size_t M, maxM, maxM0, efc;
    double mult;
    std::vector<uint32_t> l0;                  // n * (1 + maxM0): count, ids
    std::vector<std::vector<uint32_t>> up;     // per node: level * (1 + maxM)
    std::vector<int> levels;
    std::vector<std::mutex> locks;
    std::mutex global;
    int maxlevel = -1;
    int64_t entry = -1;

    Builder(const float *x_, int64_t n_, int dim_, int metric_, size_t M_, size_t efc_,
            uint64_t seed, bool exact_, bool with_l0 = true)
        : x(x_), n(n_), dim(dim_), metric(metric_), exact(exact_), M(M_), maxM(M_),
          maxM0(2 * M_), efc(std::max(efc_, M_)), l0(with_l0 ? (size_t)n_ * (1 + 2 * M_) : 0, 0),
          up((size_t)n_), levels((size_t)n_), locks((size_t)n_)
    {
        mult = 1.0 / std::log(1.0 * M);
        std::default_random_engine gen;
        gen.seed(seed);
        for (int64_t i = 0; i < n; i++) {
            std::uniform_real_distribution<double> distribution(0.0, 1.0);
            double r = -std::log(distribution(gen)) * mult;
            levels[(size_t)i] = (int)r;
        }
    }

    inline const float *vec(uint32_t i) const { return x + (size_t)i * dim; }
    inline float dist(const float *a, const float *b) const
    {
        return exact ? dist_exact(a, b, dim, metric) : dist_fast(a, b, dim, metric);
    }
    inline uint32_t *list(uint32_t i, int level)
    {
        if (level == 0) return &l0[(size_t)i * (1 + maxM0)];
        return &up[i][(size_t)(level - 1) * (1 + maxM)];
    }
    std::vector<uint32_t> copy_list(uint32_t i, int level)
    {
        std::lock_guard<std::mutex> lk(locks[i]);
        uint32_t *l = list(i, level);
        return std::vector<uint32_t>(l + 1, l + 1 + l[0]);
    }

    // HNSWLIB/hnswalg.h:153-236
    pq_t search_layer(uint32_t ep, const float *q, int layer, std::vector<uint32_t> &tag, uint32_t &epoch)
    {
        epoch++;
        if (epoch == 0) {
            std::fill(tag.begin(), tag.end(), 0);
            epoch = 1;
        }
        pq_t top, cand;
        float d = dist(q, vec(ep));
        top.emplace(d, ep);
        float lower = d;
        cand.emplace(-d, ep);
        tag[ep] = epoch;
        std::vector<uint32_t> nb;
        while (!cand.empty()) {
            pf cur = cand.top();
            if ((-cur.first) > lower && top.size() == efc) break;
            cand.pop();
            nb = copy_list(cur.second, layer);
            for (uint32_t c : nb) {
                if (tag[c] == epoch) continue;
                tag[c] = epoch;
                float d1 = dist(q, vec(c));
                if (top.size() < efc || lower > d1) {
                    cand.emplace(-d1, c);
                    top.emplace(d1, c);
                    if (top.size() > efc) top.pop();
                    if (!top.empty()) lower = top.top().first;
                }
            }
        }
        return top;
    }

    // HNSWLIB/hnswalg.h:328-369
    void heuristic(pq_t &top, size_t m)
    {
        if (top.size() < m) return;
        std::priority_queue<pf> closest;
        std::vector<pf> ret;
        while (!top.empty()) {
            closest.emplace(-top.top().first, top.top().second);
            top.pop();
        }
        while (!closest.empty()) {
            if (ret.size() >= m) break;
            pf cur = closest.top();
            float dq = -cur.first;
            closest.pop();
            bool good = true;
            for (const pf &s : ret) {
                float dd = dist(vec(s.second), vec(cur.second));
                if (dd < dq) {
                    good = false;
                    break;
                }
            }
            if (good) ret.push_back(cur);
        }
        for (const pf &c : ret) top.emplace(-c.first, c.second);
    }

    // HNSWLIB/hnswalg.h:386-501 (isUpdate = false)
    uint32_t connect(uint32_t cur, pq_t &top, int level)
    {
        size_t mcur = level ? maxM : maxM0;
        heuristic(top, M);
        std::vector<uint32_t> sel;
        sel.reserve(M);
        while (!top.empty()) {
            sel.push_back(top.top().second);
            top.pop();
        }
        uint32_t next = sel.back();
        {
            std::lock_guard<std::mutex> lk(locks[cur]);
            uint32_t *l = list(cur, level);
            l[0] = (uint32_t)sel.size();
            for (size_t i = 0; i < sel.size(); i++) l[1 + i] = sel[i];
        }
        for (uint32_t s : sel) {
            std::lock_guard<std::mutex> lk(locks[s]);
            uint32_t *l = list(s, level);
            size_t sz = l[0];
            if (sz < mcur) {
                l[1 + sz] = cur;
                l[0] = (uint32_t)(sz + 1);
            } else {
                float dmax = dist(vec(cur), vec(s));
                pq_t cands;
                cands.emplace(dmax, cur);
                for (size_t j = 0; j < sz; j++) cands.emplace(dist(vec(l[1 + j]), vec(s)), l[1 + j]);
                heuristic(cands, mcur);
                size_t idx = 0;
                while (!cands.empty()) {
                    l[1 + idx] = cands.top().second;
                    cands.pop();
                    idx++;
                }
                l[0] = (uint32_t)idx;
            }
        }
        return next;
    }

    // HNSWLIB/hnswalg.h:982-1112
    void insert(uint32_t cur, std::vector<uint32_t> &tag, uint32_t &epoch)
    {
        int curlevel = levels[cur];
        if (curlevel > 0) up[cur].assign((size_t)curlevel * (1 + maxM), 0);
        std::unique_lock<std::mutex> glk(global);
        int maxlevelcopy = maxlevel;
        int64_t ep = entry;
        if (curlevel <= maxlevelcopy) glk.unlock();
        if (ep < 0) {  // first element
            entry = cur;
            maxlevel = curlevel;
            return;
        }
        const float *q = vec(cur);
        uint32_t obj = (uint32_t)ep;
        if (curlevel < maxlevelcopy) {
            float curdist = dist(q, vec(obj));
            for (int level = maxlevelcopy; level > curlevel; level--) {
                bool changed = true;
                while (changed) {
                    changed = false;
                    std::vector<uint32_t> nb = copy_list(obj, level);
                    for (uint32_t c : nb) {
                        float d = dist(q, vec(c));
                        if (d < curdist) {
                            curdist = d;
                            obj = c;
                            changed = true;
                        }
                    }
                }
            }
        }
        for (int level = std::min(curlevel, maxlevelcopy); level >= 0; level--) {
            pq_t top = search_layer(obj, q, level, tag, epoch);
            obj = connect(cur, top, level);
        }
        if (curlevel > maxlevelcopy) {
            entry = cur;
            maxlevel = curlevel;
        }
    }
};

// Batched insertion of the levels >= 1 (the GPU builder runs layer 0).  A
// batch [a, b) follows hnswlib's addPoint (HNSWLIB/hnswalg.h:982-1112) for
// every node in it against the graph as it stood at the batch start: the
// greedy descent, then per level the construction search and the heuristic
// (M neighbours); the nodes' own lists are then written, farthest first as
// mutuallyConnectNewElement writes them (:386-440), and the reverse links
// grouped by (level, target): a target list with room for all of them takes
// them in node order (:466-470), otherwise it is re-pruned with the
// heuristic over its links and all the new ones (:471-497 with every new
// link at once).  The entry point moves to the first node of the batch with
// the highest new level.  A batch of one node is addPoint exactly.
struct UpperBatch {
    Builder b;
    std::vector<int64_t> upper_off;  // node -> first row of its levels >= 1 (or -1)
    std::vector<int32_t> flat;       // rows x maxM, -1 padded (the device layout)
    int64_t rows = 0;
    int final_maxlevel = 0;
    int64_t final_entry = 0;

    bool with_l0;  // the host runs layer 0 too (the CPU form of the batched build)

    UpperBatch(const float *x, int64_t n, int dim, int metric, size_t M, size_t efc, uint64_t seed, bool exact,
               bool with_l0_ = false)
        : b(x, n, dim, metric, M, efc, seed, exact, with_l0_), upper_off((size_t)n, -1), with_l0(with_l0_)
    {
        for (int64_t i = 0; i < n; i++) {
            const int L = b.levels[(size_t)i];
            if (L > 0) {
                upper_off[(size_t)i] = rows;
                rows += L;
                b.up[(size_t)i].assign((size_t)L * (1 + b.maxM), 0);
            }
            if (i == 0 || L > final_maxlevel) {  // addPoint moves the entry to a node above the top level
                final_entry = i;
                final_maxlevel = L;
            }
        }
        flat.assign((size_t)rows * b.maxM, -1);
    }

    void write_flat(uint32_t v, int level)
    {
        const uint32_t *l = b.list(v, level);
        int32_t *dst = &flat[(size_t)(upper_off[v] + level - 1) * b.maxM];
        for (size_t j = 0; j < b.maxM; j++) dst[j] = j < l[0] ? (int32_t)l[1 + j] : -1;
    }

    // One batch; l0_entry[i - a] receives the layer-0 entry of node i when
    // the upper levels give it one (its level-1 search), else -1 (descend).
    // Returns true when an upper list changed.
    std::vector<std::vector<uint32_t>> tags;  // per worker thread: visited tags (kept across batches)
    std::vector<uint32_t> epochs;

    bool batch(int64_t a, int64_t bnd, int nthreads, std::vector<int32_t> &l0_entry)
    {
        const int64_t a0 = a;
        l0_entry.assign((size_t)(bnd - a), -1);
        if (b.entry < 0) {  // the first node becomes the entry point (addPoint's first element)
            b.entry = a;
            b.maxlevel = b.levels[(size_t)a];
            a++;
            if (a >= bnd) return false;
        }
        const int maxlevel0 = b.maxlevel;
        const uint32_t entry0 = (uint32_t)b.entry;
        std::vector<int64_t> U;
        for (int64_t i = a; i < bnd; i++)
            if (b.levels[(size_t)i] > 0) U.push_back(i);
        // per upper node and level >= 1: the selected neighbours, closest first
        std::vector<std::vector<std::vector<pf>>> sel(U.size());
        const int nt = std::max(1, std::min(nthreads, (int)U.size()));
        if ((int)tags.size() < nt) {
            tags.resize((size_t)nt);
            epochs.resize((size_t)nt, 0);
        }
        std::atomic<size_t> next(0);
        auto work = [&](int w) {
            std::vector<uint32_t> &tag = tags[(size_t)w];
            if (tag.empty()) tag.assign((size_t)b.n, 0);
            uint32_t &epoch = epochs[(size_t)w];
            for (;;) {
                const size_t t = next.fetch_add(1);
                if (t >= U.size()) break;
                const uint32_t cur = (uint32_t)U[t];
                const int L = b.levels[cur];
                const float *q = b.vec(cur);
                uint32_t obj = entry0;
                float curdist = b.dist(q, b.vec(obj));
                for (int level = maxlevel0; level > L; level--) {
                    bool changed = true;
                    while (changed) {
                        changed = false;
                        std::vector<uint32_t> nb = b.copy_list(obj, level);
                        for (uint32_t c : nb) {
                            float d = b.dist(q, b.vec(c));
                            if (d < curdist) {
                                curdist = d;
                                obj = c;
                                changed = true;
                            }
                        }
                    }
                }
                sel[t].resize((size_t)L + 1);
                for (int level = std::min(L, maxlevel0); level >= 1; level--) {
                    pq_t top = b.search_layer(obj, q, level, tag, epoch);
                    b.heuristic(top, b.M);
                    std::vector<pf> s;
                    while (!top.empty()) {
                        s.push_back(top.top());
                        top.pop();
                    }
                    std::reverse(s.begin(), s.end());  // closest first
                    obj = s.front().second;
                    sel[t][(size_t)level] = std::move(s);
                }
                l0_entry[(size_t)(cur - a0)] = (int32_t)obj;
            }
        };
        if (U.empty()) {
        } else if (nt <= 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++) th.emplace_back(work, t);
            for (auto &t : th) t.join();
        }
        // own lists, farthest first
        struct Rev { int level; uint32_t s, u; };
        std::vector<Rev> rev;
        for (size_t t = 0; t < U.size(); t++) {
            const uint32_t cur = (uint32_t)U[t];
            for (int level = 1; level < (int)sel[t].size(); level++) {
                const std::vector<pf> &s = sel[t][(size_t)level];
                if (s.empty()) continue;
                uint32_t *l = b.list(cur, level);
                l[0] = (uint32_t)s.size();
                for (size_t i = 0; i < s.size(); i++) l[1 + i] = s[s.size() - 1 - i].second;
                write_flat(cur, level);
                for (const pf &e : s) rev.push_back({level, e.second, cur});
            }
        }
        std::sort(rev.begin(), rev.end(), [](const Rev &x, const Rev &y) {
            return x.level != y.level ? x.level < y.level : x.s != y.s ? x.s < y.s : x.u < y.u;
        });
        for (size_t i = 0; i < rev.size();) {
            size_t j = i;
            while (j < rev.size() && rev[j].level == rev[i].level && rev[j].s == rev[i].s) j++;
            const int level = rev[i].level;
            const uint32_t s = rev[i].s;
            uint32_t *l = b.list(s, level);
            const size_t sz = l[0];
            if (sz + (j - i) <= b.maxM) {
                for (size_t t = i; t < j; t++) l[1 + sz + (t - i)] = rev[t].u;
                l[0] = (uint32_t)(sz + (j - i));
            } else {
                pq_t cands;
                for (size_t t = i; t < j; t++) cands.emplace(b.dist(b.vec(rev[t].u), b.vec(s)), rev[t].u);
                for (size_t t = 0; t < sz; t++) cands.emplace(b.dist(b.vec(l[1 + t]), b.vec(s)), l[1 + t]);
                b.heuristic(cands, b.maxM);
                size_t idx = 0;
                while (!cands.empty()) {
                    l[1 + idx] = cands.top().second;
                    cands.pop();
                    idx++;
                }
                l[0] = (uint32_t)idx;
            }
            write_flat(s, level);
            i = j;
        }
        // the entry point: the first node of the batch with the highest new level
        for (int64_t v : U)
            if (b.levels[(size_t)v] > b.maxlevel) {
                b.maxlevel = b.levels[(size_t)v];
                b.entry = v;
            }
        if (with_l0) layer0(a, a0, bnd, nthreads, l0_entry);
        return !U.empty();
    }

    // Layer 0 of the batch [a, bnd) on the host, as the GPU builder runs it:
    // entry from the levels >= 1 (or the greedy descent from the entry point
    // over the levels as they now stand), construction search over layer 0
    // as it stood at the batch start, heuristic (M), own rows, reverse links
    // grouped by target (maxM0).
    void layer0(int64_t a, int64_t a0, int64_t bnd, int nthreads, const std::vector<int32_t> &l0_entry)
    {
        if (a >= bnd) return;
        const int64_t nb = bnd - a;
        std::vector<std::vector<pf>> sel((size_t)nb);
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, nb));
        if ((int)tags.size() < nt) {
            tags.resize((size_t)nt);
            epochs.resize((size_t)nt, 0);
        }
        std::atomic<int64_t> next(0);
        const uint32_t entry = (uint32_t)b.entry;
        const int maxlevel = b.maxlevel;
        auto work = [&](int w) {
            std::vector<uint32_t> &tag = tags[(size_t)w];
            if (tag.empty()) tag.assign((size_t)b.n, 0);
            uint32_t &epoch = epochs[(size_t)w];
            for (;;) {
                const int64_t t = next.fetch_add(1);
                if (t >= nb) break;
                const uint32_t cur = (uint32_t)(a + t);
                const float *q = b.vec(cur);
                uint32_t obj;
                if (l0_entry[(size_t)(cur - a0)] >= 0) {
                    obj = (uint32_t)l0_entry[(size_t)(cur - a0)];
                } else {
                    obj = entry;
                    float curdist = b.dist(q, b.vec(obj));
                    for (int level = maxlevel; level > 0; level--) {
                        bool changed = true;
                        while (changed) {
                            changed = false;
                            std::vector<uint32_t> nbl = b.copy_list(obj, level);
                            for (uint32_t c : nbl) {
                                float d = b.dist(q, b.vec(c));
                                if (d < curdist) {
                                    curdist = d;
                                    obj = c;
                                    changed = true;
                                }
                            }
                        }
                    }
                }
                pq_t top = b.search_layer(obj, q, 0, tag, epoch);
                b.heuristic(top, b.M);
                std::vector<pf> sv;
                while (!top.empty()) {
                    sv.push_back(top.top());
                    top.pop();
                }
                sel[(size_t)t] = std::move(sv);  // farthest first
            }
        };
        if (nt <= 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++) th.emplace_back(work, t);
            for (auto &t : th) t.join();
        }
        std::vector<std::pair<uint32_t, uint32_t>> rev;  // (target, new node)
        for (int64_t t = 0; t < nb; t++) {
            const uint32_t cur = (uint32_t)(a + t);
            uint32_t *l = b.list(cur, 0);
            const std::vector<pf> &sv = sel[(size_t)t];
            l[0] = (uint32_t)sv.size();
            for (size_t i = 0; i < sv.size(); i++) {
                l[1 + i] = sv[i].second;
                rev.emplace_back(sv[i].second, cur);
            }
        }
        std::sort(rev.begin(), rev.end());
        for (size_t i = 0; i < rev.size();) {
            size_t j = i;
            while (j < rev.size() && rev[j].first == rev[i].first) j++;
            const uint32_t s = rev[i].first;
            uint32_t *l = b.list(s, 0);
            // as k_build_link: the new links in chunks of at most 1024 - maxM0
            const size_t chunk = 1024 - b.maxM0;
            for (size_t t0 = i; t0 < j; t0 += chunk) {
                const size_t t1 = std::min(j, t0 + chunk);
                const size_t sz = l[0];
                if (sz + (t1 - t0) <= b.maxM0) {
                    for (size_t t = t0; t < t1; t++) l[1 + sz + (t - t0)] = rev[t].second;
                    l[0] = (uint32_t)(sz + (t1 - t0));
                    continue;
                }
                pq_t cands;
                for (size_t t = t0; t < t1; t++) cands.emplace(b.dist(b.vec(rev[t].second), b.vec(s)), rev[t].second);
                for (size_t t = 0; t < sz; t++) cands.emplace(b.dist(b.vec(l[1 + t]), b.vec(s)), l[1 + t]);
                b.heuristic(cands, b.maxM0);
                size_t idx = 0;
                while (!cands.empty()) {
                    l[1 + idx] = cands.top().second;
                    cands.pop();
                    idx++;
                }
                l[0] = (uint32_t)idx;
            }
            i = j;
        }
    }
};

}  // namespace

namespace gvs {
// Host side of the GPU builder (levels >= 1); see UpperBatch.
struct UpperBuild {
    UpperBatch u;
    UpperBuild(const float *x, int64_t n, int dim, int metric, int M, int efc, uint64_t seed, bool exact)
        : u(x, n, dim, metric, (size_t)M, (size_t)efc, seed, exact)
    {
    }
};

UpperBuild *upper_build_create(const float *x, int64_t n, int dim, int metric, int M, int efc, uint64_t seed,
                               bool exact)
{
    return new UpperBuild(x, n, dim, metric, M, efc, seed, exact);
}
void upper_build_free(UpperBuild *ub) { delete ub; }
bool upper_build_batch(UpperBuild *ub, int64_t a, int64_t b, int nthreads, std::vector<int32_t> &l0_entry)
{
    return ub->u.batch(a, b, nthreads, l0_entry);
}
int upper_build_entry(const UpperBuild *ub) { return (int)ub->u.b.entry; }
int upper_build_maxlevel(const UpperBuild *ub) { return ub->u.b.maxlevel; }
int upper_build_final_maxlevel(const UpperBuild *ub) { return ub->u.final_maxlevel; }
int64_t upper_build_final_entry(const UpperBuild *ub) { return ub->u.final_entry; }
const std::vector<int64_t> &upper_build_offsets(const UpperBuild *ub) { return ub->u.upper_off; }
const std::vector<int32_t> &upper_build_rows(const UpperBuild *ub) { return ub->u.flat; }
int64_t upper_build_nrows(const UpperBuild *ub) { return ub->u.rows; }
const std::vector<int> &upper_build_levels(const UpperBuild *ub) { return ub->u.b.levels; }
int upper_build_efc(const UpperBuild *ub) { return (int)ub->u.b.efc; }
double upper_build_mult(const UpperBuild *ub) { return ub->u.b.mult; }
}  // namespace gvs

// The batched insertion entirely on the host: the rules of
// gvs_hnsw_build_device with the same batch schedule (a check of the GPU
// builder: with exact_order both give the same graph).
extern "C" int gvs_hnsw_build_batched(const float *x, int64_t n, int32_t dim, int32_t metric, int32_t M,
                                      int32_t ef_construction, uint64_t seed, int32_t nthreads, int32_t exact_order,
                                      double batch_ratio, int32_t max_batch, gvs_host_graph **out)
{
    if (!x || n <= 0 || n > INT32_MAX || dim <= 0 || M < 2 || M > 1024 || ef_construction < 1 || !out ||
        (metric != GVS_METRIC_L2 && metric != GVS_METRIC_IP)) {
        gvs::set_error("gvs_hnsw_build_batched: invalid argument");
        return GVS_ERR_ARG;
    }
    try {
        UpperBatch ub(x, n, dim, metric, (size_t)M, (size_t)ef_construction, seed, exact_order != 0, true);
        const double ratio = batch_ratio > 0.0 ? batch_ratio : (batch_ratio < 0.0 ? 0.0 : 0.01);
        const int64_t mb = max_batch > 0 ? max_batch : 32768;
        std::vector<int32_t> l0e;
        for (int64_t a = 0; a < n;) {
            const int64_t size = a == 0 ? 1 : std::max<int64_t>(1, std::min<int64_t>(mb, (int64_t)(a * ratio)));
            const int64_t bnd = std::min(n, a + size);
            ub.batch(a, bnd, std::max(1, (int)nthreads), l0e);
            a = bnd;
        }
        Builder &b = ub.b;
        gvs_host_graph *g = new gvs_host_graph();
        g->n = n;
        g->dim = dim;
        g->metric = metric;
        g->vectors.assign(x, x + (size_t)n * dim);
        g->r0 = (int32_t)b.maxM0;
        g->adj0.assign((size_t)n * b.maxM0, -1);
        for (int64_t i = 0; i < n; i++) {
            const uint32_t *l = &b.l0[(size_t)i * (1 + b.maxM0)];
            for (uint32_t j = 0; j < l[0]; j++) g->adj0[(size_t)i * b.maxM0 + j] = (int32_t)l[1 + j];
        }
        g->max_level = b.maxlevel;
        g->entry = (int32_t)b.entry;
        g->ru = (int32_t)b.maxM;
        g->levels.assign(b.levels.begin(), b.levels.end());
        g->upper_off = ub.upper_off;
        g->adj_up = ub.flat;
        g->labels.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = i;
        g->M = M;
        g->ef_construction = (int32_t)b.efc;
        g->mult = b.mult;
        *out = g;
        return GVS_OK;
    } catch (const std::bad_alloc &) {
        gvs::set_error("gvs_hnsw_build_batched: out of host memory");
        return GVS_ERR_NOMEM;
    }
}

extern "C" int gvs_hnsw_build(const float *x, int64_t n, int32_t dim, int32_t metric, int32_t M,
                              int32_t ef_construction, uint64_t seed, int32_t nthreads,
                              int32_t exact_order, gvs_host_graph **out)
{
    if (!x || n <= 0 || dim <= 0 || M < 2 || ef_construction < 1 || !out || n > INT32_MAX ||
        (metric != GVS_METRIC_L2 && metric != GVS_METRIC_IP)) {
        gvs::set_error("gvs_hnsw_build: invalid argument");
        return GVS_ERR_ARG;
    }
    if (M > 1024) {
        gvs::set_error("gvs_hnsw_build: M too large");
        return GVS_ERR_ARG;
    }
    try {
        Builder b(x, n, dim, metric, (size_t)M, (size_t)ef_construction, seed, exact_order != 0);
        int nt = std::max(1, (int)nthreads);
        {
            std::vector<uint32_t> tag((size_t)n, 0);
            uint32_t epoch = 0;
            b.insert(0, tag, epoch);
        }
        std::atomic<int64_t> next(1);
        // GVS_VERBOSE=1: one stderr line per tenth of the points inserted
        const char *verbose = std::getenv("GVS_VERBOSE");
        const int64_t report = (verbose && verbose[0] == '1') ? std::max<int64_t>(1, n / 10) : 0;
        auto work = [&]() {
            std::vector<uint32_t> tag((size_t)n, 0);
            uint32_t epoch = 0;
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= n) break;
                b.insert((uint32_t)i, tag, epoch);
                if (report && (i + 1) % report == 0)
                    std::fprintf(stderr, "gvs_hnsw_build: %lld of %lld points inserted\n", (long long)(i + 1),
                                 (long long)n);
            }
        };
        if (nt == 1) {
            work();
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++) th.emplace_back(work);
            for (auto &t : th) t.join();
        }

        gvs_host_graph *g = new gvs_host_graph();
        g->n = n;
        g->dim = dim;
        g->metric = metric;
        g->vectors.assign(x, x + (size_t)n * dim);
        g->r0 = (int32_t)b.maxM0;
        g->adj0.assign((size_t)n * b.maxM0, -1);
        for (int64_t i = 0; i < n; i++) {
            const uint32_t *l = &b.l0[(size_t)i * (1 + b.maxM0)];
            for (uint32_t j = 0; j < l[0]; j++) g->adj0[(size_t)i * b.maxM0 + j] = (int32_t)l[1 + j];
        }
        g->max_level = b.maxlevel;
        g->entry = (int32_t)b.entry;
        g->ru = (int32_t)b.maxM;
        g->levels.assign(b.levels.begin(), b.levels.end());
        g->upper_off.assign((size_t)n, -1);
        int64_t rows = 0;
        for (int64_t i = 0; i < n; i++)
            if (b.levels[(size_t)i] > 0) {
                g->upper_off[(size_t)i] = rows;
                rows += b.levels[(size_t)i];
            }
        g->adj_up.assign((size_t)rows * b.maxM, -1);
        for (int64_t i = 0; i < n; i++) {
            int L = b.levels[(size_t)i];
            for (int l = 1; l <= L; l++) {
                const uint32_t *src = &b.up[(size_t)i][(size_t)(l - 1) * (1 + b.maxM)];
                int32_t *dst = &g->adj_up[(size_t)(g->upper_off[(size_t)i] + l - 1) * b.maxM];
                for (uint32_t j = 0; j < src[0]; j++) dst[j] = (int32_t)src[1 + j];
            }
        }
        g->labels.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) g->labels[(size_t)i] = i;
        g->M = M;
        g->ef_construction = (int32_t)b.efc;
        g->mult = b.mult;
        *out = g;
        return GVS_OK;
    } catch (const std::bad_alloc &) {
        gvs::set_error("gvs_hnsw_build: out of host memory");
        return GVS_ERR_NOMEM;
    }
}

// Flat (NSG-style) graph from the HNSW base layer: rows pruned to <= R with
// the occlusion rule of getNeighborsByHeuristic2, entry = navigating node
// (the node closest to the centroid of the data), no upper layers.
extern "C" int gvs_nsg_from_hnsw(const gvs_host_graph *from, int32_t R, gvs_host_graph **out)
{
    if (!from || !out || R < 1 || R > 1024) {
        gvs::set_error("gvs_nsg_from_hnsw: invalid argument");
        return GVS_ERR_ARG;
    }
    const int64_t n = from->n;
    const int dim = from->dim;
    gvs_host_graph *g = new gvs_host_graph();
    g->n = n;
    g->dim = dim;
    g->metric = from->metric;
    g->vectors = from->vectors;
    g->r0 = R;
    g->adj0.assign((size_t)n * R, -1);
    auto vec = [&](int64_t i) { return from->vectors.data() + (size_t)i * dim; };
    const bool hier = from->max_level > 0 && !from->upper_off.empty() && from->levels.size() == (size_t)n;
    for (int64_t i = 0; i < n; i++) {
        const int32_t *row = &from->adj0[(size_t)i * from->r0];
        int cnt = 0;
        while (cnt < from->r0 && row[cnt] >= 0) cnt++;
        // Long-range edges: a node of level L >= 1 keeps its upper-layer
        // links (top level first, at most R/2).  A flat graph searched from
        // one navigating node needs them: layer 0 alone only links near
        // neighbours, so greedy search from one entry stalls between
        // clusters.  (NSG gets the same property from its construction.)
        std::vector<int32_t> keep;
        auto has = [&](int32_t u) { return std::find(keep.begin(), keep.end(), u) != keep.end(); };
        if (hier && from->levels[(size_t)i] > 0) {
            for (int l = from->levels[(size_t)i]; l >= 1 && (int)keep.size() < R / 2; l--) {
                const int32_t *urow = &from->adj_up[(size_t)(from->upper_off[(size_t)i] + l - 1) * from->ru];
                for (int j = 0; j < from->ru && (int)keep.size() < R / 2; j++)
                    if (urow[j] >= 0 && urow[j] != (int32_t)i && !has(urow[j])) keep.push_back(urow[j]);
            }
        }
        std::vector<int32_t> rest;
        for (int j = 0; j < cnt; j++)
            if (!has(row[j])) rest.push_back(row[j]);
        if ((int)(keep.size() + rest.size()) <= R) {
            keep.insert(keep.end(), rest.begin(), rest.end());
        } else {
            // occlusion rule over the layer-0 candidates (closest first): a
            // candidate closer to an already kept neighbour than to the node
            // is skipped; then fill up to R with the nearest skipped ones
            std::vector<std::pair<float, int32_t>> c;
            for (int32_t u : rest) c.emplace_back(dist_fast(vec(i), vec(u), dim, from->metric), u);
            std::sort(c.begin(), c.end());
            const size_t base = keep.size();
            for (auto &p : c) {
                if ((int)keep.size() >= R) break;
                bool good = true;
                for (size_t t = base; t < keep.size(); t++)
                    if (dist_fast(vec(keep[t]), vec(p.second), dim, from->metric) < p.first) {
                        good = false;
                        break;
                    }
                if (good) keep.push_back(p.second);
            }
            for (auto &p : c) {
                if ((int)keep.size() >= R) break;
                if (!has(p.second)) keep.push_back(p.second);
            }
        }
        for (size_t j = 0; j < keep.size(); j++) g->adj0[(size_t)i * R + j] = keep[j];
    }
    // navigating node
    std::vector<double> cen((size_t)dim, 0.0);
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) cen[(size_t)d] += vec(i)[d];
    std::vector<float> cf((size_t)dim);
    for (int d = 0; d < dim; d++) cf[(size_t)d] = (float)(cen[(size_t)d] / (double)n);
    // the navigating node: the node nearest to the centroid, among the nodes
    // with long-range edges when the source graph has upper layers
    float best = INFINITY;
    int32_t bi = 0;
    for (int64_t i = 0; i < n; i++) {
        if (hier && from->levels[(size_t)i] == 0) continue;
        float d = dist_fast(cf.data(), vec(i), dim, GVS_METRIC_L2);
        if (d < best) {
            best = d;
            bi = (int32_t)i;
        }
    }
    g->max_level = 0;
    g->entry = bi;
    g->ru = 0;
    g->levels.assign((size_t)n, 0);
    g->labels = from->labels;
    *out = g;
    return GVS_OK;
}
