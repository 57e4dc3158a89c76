// Seeded Gaussian-mixture generator for the synthetic workloads (the
// reference's datasets - SIFT, Deep, SBERT, SPACEV - are not available).
// Counter-based: every centre and every point draws from its own
// splitmix64 stream keyed by (seed, index), so the output does not depend on
// the thread count.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "host_graph.hpp"

namespace {
struct Stream {
    uint64_t s;
    explicit Stream(uint64_t seed, uint64_t index, uint64_t salt)
        : s(seed ^ (0x632BE59BD9B4E019ull * (index + 1)) ^ salt)
    {
        next();
    }
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    // Box-Muller, both outputs used
    bool have = false;
    double spare = 0.0;
    double normal()
    {
        if (have) {
            have = false;
            return spare;
        }
        double u1 = u01(), u2 = u01();
        double r = std::sqrt(-2.0 * std::log(u1));
        double t = 6.283185307179586476925286766559 * u2;
        spare = r * std::sin(t);
        have = true;
        return r * std::cos(t);
    }
};
}  // namespace

extern "C" int gvs_gen_mixture(uint64_t seed_centers, uint64_t seed_points, int64_t n, int32_t dim,
                               int32_t ncenters, float center_std, float noise_std, int32_t normalize,
                               float *out, int32_t nthreads)
{
    if (!out || n < 0 || dim <= 0 || ncenters <= 0) {
        gvs::set_error("gvs_gen_mixture: invalid argument");
        return GVS_ERR_ARG;
    }
    std::vector<double> cen((size_t)ncenters * dim);
    for (int32_t c = 0; c < ncenters; c++) {
        Stream st(seed_centers, (uint64_t)c, 0xC3A5C85C97CB3127ull);
        for (int32_t d = 0; d < dim; d++) cen[(size_t)c * dim + d] = center_std * st.normal();
    }
    int nt = std::max(1, (int)nthreads);
    auto work = [&](int t) {
        for (int64_t i = t; i < n; i += nt) {
            Stream st(seed_points, (uint64_t)i, 0x9AE16A3B2F90404Full);
            uint32_t u = (uint32_t)(st.next() % (uint64_t)ncenters);
            float *row = out + (size_t)i * dim;
            double nrm = 0.0;
            for (int32_t d = 0; d < dim; d++) {
                double v = cen[(size_t)u * dim + d] + noise_std * st.normal();
                row[d] = (float)v;
                nrm += (double)row[d] * row[d];
            }
            if (normalize && nrm > 0.0) {
                float s = (float)(1.0 / std::sqrt(nrm));
                for (int32_t d = 0; d < dim; d++) row[d] *= s;
            }
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto &t : th) t.join();
    }
    return GVS_OK;
}
