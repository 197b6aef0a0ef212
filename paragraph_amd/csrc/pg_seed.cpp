// pg_seed.cpp -- synthetic seed graphs and log-fitness draws (host).
//
// The reference's seed graphs ("sj" edgelist/nodelist, src/graph.cu:336-447)
// are not available, so inputs are generated: birth years uniform in
// [year_lo, year_hi], paper ids sorted by birth (the reference numbers
// papers densely, src/graph.cu:361-370), and each paper cites
// min(Poisson(mean), #older) distinct older papers uniformly.  Log-fitness
// is Normal(mu, sigma^2) (lognormal fitness), replacing the reference's
// power-law fitness table (src/abm.cu:124-176).  Every draw is a Philox
// output keyed by the seed, so any machine produces the same graph.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_math.h"

using namespace pg;

namespace {

template <typename F>
void parallel_for(int64_t n, F&& f) {
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nt = std::max<int64_t>(1, std::min<int64_t>(hw ? hw : 1, 16));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (int64_t t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &f] { for (int64_t i = b; i < e; ++i) f(i); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int pg_seed_graph(uint64_t seed, int64_t n, int32_t year_lo, int32_t year_hi,
                             double mean_cites, int32_t* birth, int64_t* n_edges, int32_t* src,
                             int32_t* dst) {
    if (n < 1 || n >= (int64_t(1) << 31) - 1 || year_hi < year_lo || !(mean_cites >= 0.0) ||
        !birth || !n_edges)
        return PG_EINVAL;
    const int64_t span = (int64_t)year_hi - year_lo + 1;
    // birth years drawn per draw index, then ids sorted by year (stable)
    std::vector<int32_t> raw(n);
    parallel_for(n, [&](int64_t j) {
        raw[j] = (int32_t)(year_lo + seed_birth_offset(seed, (uint32_t)j, span));
    });
    std::vector<int64_t> per_year(span + 1, 0);
    for (int64_t j = 0; j < n; ++j) per_year[raw[j] - year_lo + 1]++;
    for (int64_t y = 0; y < span; ++y) per_year[y + 1] += per_year[y];
    for (int64_t y = 0; y < span; ++y)
        for (int64_t i = per_year[y]; i < per_year[y + 1]; ++i) birth[i] = (int32_t)(year_lo + y);
    // citation counts
    std::vector<int32_t> cnt(n);
    const double e0 = std::exp(-mean_cites);
    parallel_for(n, [&](int64_t i) {
        const int64_t older = per_year[birth[i] - year_lo];
        cnt[i] = seed_count(seed, (uint32_t)i, mean_cites, e0, older);
    });
    std::vector<int64_t> off(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
    if (!src) { *n_edges = off[n]; return PG_OK; }
    if (!dst || *n_edges < off[n]) return PG_EINVAL;
    *n_edges = off[n];
    parallel_for(n, [&](int64_t i) {
        const int64_t older = per_year[birth[i] - year_lo];
        const int32_t c = cnt[i];
        int32_t* out = dst + off[i];
        int32_t got = 0;
        for (uint32_t a = 0; got < c; ++a) {
            const int32_t t = seed_target(seed, (uint32_t)i, a, older);
            bool dup = false;
            for (int32_t q = 0; q < got && !dup; ++q) dup = (out[q] == (int32_t)t);
            if (!dup) out[got++] = (int32_t)t;
        }
        for (int32_t q = 0; q < c; ++q) src[off[i] + q] = (int32_t)i;
    });
    return PG_OK;
}

extern "C" int pg_seed_logfit(uint64_t seed, int64_t first, int64_t count, double mu, double sigma,
                              float* out) {
    if (first < 0 || count < 0 || (count > 0 && !out) || !(sigma >= 0.0)) return PG_EINVAL;
    parallel_for(count, [&](int64_t k) {
        out[k] = seed_logfit(seed, (uint32_t)(first + k), mu, sigma);
    });
    return PG_OK;
}
