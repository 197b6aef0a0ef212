// pg_math.h -- counter-based RNG and deterministic elementary functions used
// by every kernel and by the host-side seed generator.
//
// Philox4x32-10 follows Salmon et al., "Parallel random numbers: as easy as
// 1, 2, 3" (SC'11).  It replaces the reference's curand XORWOW streams
// (src/kernel.cu:233, src/abm.cu:529), which were seeded from the wall clock
// and so could not be reproduced.  Counter layout: (paper id, attempt,
// growth-year index, tag); key: the 64-bit simulation seed.
//
// exp/log use Cody-Waite reduction with Taylor polynomials (Horner, explicit
// fma) and the atanh series log(1+f) = 2 atanh(f/(2+f)).  Each operation is
// spelled out and the build disables FP contraction, so host and gfx950
// produce the same bits for the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG_HD __host__ __device__ __forceinline__

namespace pg {

enum : uint32_t { TAG_PA = 0, TAG_NB = 1, TAG_ER = 2, TAG_SEED_BIRTH = 3,
                  TAG_FITNESS = 4, TAG_SEED_COUNT = 5, TAG_SEED_CITE = 6,
                  // the reference's scoring model: fitness peak value, paper
                  // properties, generator, same-year citation, selection
                  // key, uniform fill
                  TAG_ABM_FIT = 7, TAG_ABM_PAPER = 8, TAG_ABM_GEN = 9, TAG_ABM_SAME = 10,
                  TAG_ABM_KEY = 11, TAG_ABM_UNI = 12 };

struct u32x4 { uint32_t x, y, z, w; };

PG_HD u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return u32x4{c0, c1, c2, c3};
}

// Uniform in [0,1) with 53 random bits.
PG_HD double u53(uint32_t x0, uint32_t x1) {
    uint64_t v = ((uint64_t)x0 << 21) ^ (uint64_t)(x1 >> 11);
    return (double)v * 0x1p-53;
}
// Uniform in (0,1].
PG_HD double u53_open0(uint32_t x0, uint32_t x1) {
    uint64_t v = ((uint64_t)x0 << 21) ^ (uint64_t)(x1 >> 11);
    return (double)(v + 1) * 0x1p-53;
}

PG_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
PG_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
PG_HD uint64_t d2u(double x) { return __builtin_bit_cast(uint64_t, x); }
PG_HD double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }

// exp(x) in FP32: Cody-Waite reduction by ln2 (hi/lo split), degree-7
// Taylor polynomial on [-ln2/2, ln2/2], scaled by 2^n.  Written without
// branches (the special cases select over the polynomial's result, which is
// then discarded): every lane runs the same instructions.  The scaling is
// one ldexp: for the n this range reaches (-126..128) p*2^n is a single
// correctly rounded scaling, the same value as scaling by two halves of n.
// Results below the smallest normal are 0; a NaN input gives a NaN.
PG_HD float det_expf(float x) {
    const float kf = __builtin_rintf(x * 0x1.715476p+0f);
    float r = __builtin_fmaf(kf, -0x1.62e4p-1f, x);
    r = __builtin_fmaf(kf, -0x1.7f7d1cp-20f, r);
    float p = 1.0f / 5040.0f;
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    // kf clamped (a NaN goes to a bound) so the conversion is defined; a NaN
    // x carries through the polynomial, out-of-range x are selected below
    const int n = (int)__builtin_fminf(__builtin_fmaxf(kf, -200.0f), 200.0f);
    float v = __builtin_ldexpf(p, n);
    v = x < -87.3365402f ? 0.0f : v;
    return x > 88.7228317f ? __builtin_huge_valf() : v;
}

PG_HD float det_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    uint32_t b = f2u(x);
    int e = 0;
    if (b < 0x00800000u) { x = x * 8388608.0f; b = f2u(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float mm = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (mm > 1.41421354f) { mm = mm * 0.5f; e += 1; }
    float f = mm - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = __builtin_fmaf(z, 2.0f / 11.0f, 2.0f / 9.0f);
    R = __builtin_fmaf(z, R, 2.0f / 7.0f);
    R = __builtin_fmaf(z, R, 2.0f / 5.0f);
    R = __builtin_fmaf(z, R, 2.0f / 3.0f);
    R = z * R;
    float t = f - R;
    t = s * t;
    float lm = f - t;
    float ef = (float)e;
    return __builtin_fmaf(ef, 0x1.62e4p-1f, __builtin_fmaf(ef, 0x1.7f7d1cp-20f, lm));
}

PG_HD double det_log_core(double f) {
    double s = f / (2.0 + f);
    double z = s * s;
    double R = __builtin_fma(z, 2.0 / 21.0, 2.0 / 19.0);
    R = __builtin_fma(z, R, 2.0 / 17.0);
    R = __builtin_fma(z, R, 2.0 / 15.0);
    R = __builtin_fma(z, R, 2.0 / 13.0);
    R = __builtin_fma(z, R, 2.0 / 11.0);
    R = __builtin_fma(z, R, 2.0 / 9.0);
    R = __builtin_fma(z, R, 2.0 / 7.0);
    R = __builtin_fma(z, R, 2.0 / 5.0);
    R = __builtin_fma(z, R, 2.0 / 3.0);
    R = z * R;
    double t = f - R;
    t = s * t;
    return f - t;
}

PG_HD double det_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_huge_val();
    if (x == __builtin_huge_val()) return x;
    uint64_t b = d2u(x);
    int e = 0;
    if (b < 0x0010000000000000ull) { x = x * 0x1p52; b = d2u(x); e = -52; }
    e += (int)(b >> 52) - 1023;
    double mm = u2d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (mm > 1.4142135623730951) { mm = mm * 0.5; e += 1; }
    double lm = det_log_core(mm - 1.0);
    double ef = (double)e;
    return __builtin_fma(ef, 6.93147180369123816490e-01,
                         __builtin_fma(ef, 1.90821492927058770002e-10, lm));
}

// det_log for a positive, normal, finite x: the same operations on that
// path, without the special cases (a partition sum minus earlier picks).
PG_HD double det_log_pos(double x) {
    const uint64_t b = d2u(x);
    int e = (int)(b >> 52) - 1023;
    double mm = u2d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    const bool hi = mm > 1.4142135623730951;
    mm = hi ? mm * 0.5 : mm;
    e += hi ? 1 : 0;
    const double lm = det_log_core(mm - 1.0);
    const double ef = (double)e;
    return __builtin_fma(ef, 6.93147180369123816490e-01,
                         __builtin_fma(ef, 1.90821492927058770002e-10, lm));
}

PG_HD double det_log1p(double x) {
    if (x > -0.29289321881345248 && x < 0.41421356237309515) return det_log_core(x);
    return det_log(1.0 + x);
}

// ---- synthetic seed graphs (host generator and device generator) ----
// Draws keyed by (paper, attempt, kSeedYear, tag) under the graph's seed.
constexpr uint32_t kSeedYear = 0xFFFFFFFFu;

// Raw birth-year offset in [0, span) of draw j.
PG_HD int64_t seed_birth_offset(uint64_t seed, uint32_t j, int64_t span) {
    const u32x4 r = philox(j, 0u, kSeedYear, TAG_SEED_BIRTH, seed);
    int64_t y = (int64_t)__builtin_floor(u53(r.x, r.y) * (double)span);
    return y >= span ? span - 1 : y;
}

// Poisson(lambda) by inversion of uniform u: the smallest k with F(k) > u;
// e0 = exp(-lambda), computed once by the caller.
PG_HD int32_t poisson_inv(double lambda, double e0, double u) {
    double p = e0, F = p;
    int32_t k = 0;
    while (u >= F && k < 100000) {
        ++k;
        p *= lambda / k;
        F += p;
        if (p == 0.0) break;
    }
    return k;
}

PG_HD int32_t seed_count(uint64_t seed, uint32_t i, double lambda, double e0, int64_t older) {
    const u32x4 r = philox(i, 0u, kSeedYear, TAG_SEED_COUNT, seed);
    const int64_t k = poisson_inv(lambda, e0, u53(r.x, r.y));
    return (int32_t)(k < older ? k : older);
}

// Attempt a of paper i's citation draws: uniform among its `older` papers.
PG_HD int32_t seed_target(uint64_t seed, uint32_t i, uint32_t a, int64_t older) {
    const u32x4 r = philox(i, a, kSeedYear, TAG_SEED_CITE, seed);
    int64_t t = (int64_t)__builtin_floor(u53(r.x, r.y) * (double)older);
    return (int32_t)(t >= older ? older - 1 : t);
}

// Log-fitness of paper g: mu + sigma * z, z standard normal by the polar
// method (Marsaglia & Bray, 1964) on successive Philox draws.
PG_HD float seed_logfit(uint64_t seed, uint32_t g, double mu, double sigma) {
    double z = 0.0;
    for (uint32_t a = 0;; ++a) {
        const u32x4 r = philox(g, a, kSeedYear, TAG_FITNESS, seed);
        const double v1 = 2.0 * u53(r.x, r.y) - 1.0;
        const double v2 = 2.0 * u53(r.z, r.w) - 1.0;
        const double s = v1 * v1 + v2 * v2;
        if (s >= 1.0 || s == 0.0) continue;
        z = v1 * __builtin_sqrt(-2.0 * det_log(s) / s);
        break;
    }
    return (float)(mu + sigma * z);
}

// ---- the reference's scoring model (agent-based model, src/abm.cu) ----
// Per-paper properties are keyed by (paper id, word, kSeedYear, tag); the
// year's draws by (citing paper, word, year index, tag).
constexpr int kAbmMaxOut = 250;   // the reference's citation buffer (src/kernel.cu:1151)
constexpr int kAbmFitMax = 1000;  // fitness peak values 1..1000 (includes/abm.cuh:543-544)

// floor(u * n) for the 53-bit uniform of (a, b), clamped to n-1.
PG_HD int64_t abm_pick(uint32_t a, uint32_t b, int64_t n) {
    const int64_t x = (int64_t)__builtin_floor(u53(a, b) * (double)n);
    return x >= n ? n - 1 : x;
}

// Static fitness peak value of paper id: the first f with cdf[f-1] > u*total
// (cdf: the 1000 cumulative weights f^-1.634, includes/abm.cuh:469-484).
PG_HD int32_t abm_fitness(uint64_t seed, uint32_t id, const double* cdf) {
    const u32x4 r = philox(id, 0u, kSeedYear, TAG_ABM_FIT, seed);
    const double t = u53(r.x, r.y) * cdf[kAbmFitMax - 1];
    int lo = 0, hi = kAbmFitMax - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1;
    }
    return lo + 1;
}

struct AbmPaper {
    int32_t k;        // out-degree drawn from the bag
    double alpha;     // neighbourhood share
    double w[3];      // normalised preferential, recency, fitness weights
};

// Out-degree, share and weights of paper id (PopulateOutDegreeArr,
// PopulateAlphaArr, PopulateWeightArrs: src/abm.cu:241-290).  alpha < 0
// draws a uniform share; any weight equal to -1 draws three uniform weights.
PG_HD AbmPaper abm_paper(uint64_t seed, uint32_t id, const int32_t* bag, int32_t n_bag,
                         double alpha, double w_pa, double w_rec, double w_fit) {
    AbmPaper p;
    const u32x4 r = philox(id, 0u, kSeedYear, TAG_ABM_PAPER, seed);
    p.k = bag[abm_pick(r.x, r.y, n_bag)];
    p.alpha = alpha >= 0.0 ? alpha : u53(r.z, r.w);
    double a = w_pa, b = w_rec, c = w_fit;
    if (w_pa == -1.0 || w_rec == -1.0 || w_fit == -1.0) {
        const u32x4 q = philox(id, 1u, kSeedYear, TAG_ABM_PAPER, seed);
        const u32x4 q2 = philox(id, 2u, kSeedYear, TAG_ABM_PAPER, seed);
        a = u53(q.x, q.y);
        b = u53(q.z, q.w);
        c = u53(q2.x, q2.y);
    }
    const double s = a + b + c;
    if (s > 0.0) {
        p.w[0] = a / s;
        p.w[1] = b / s;
        p.w[2] = c / s;
    } else {
        p.w[0] = p.w[1] = p.w[2] = 1.0 / 3.0;
    }
    return p;
}

// Citation budgets of one new paper (src/kernel.cu:244-255, :383-386):
// reserved = floor(frc*k), rem = k-1-sy-reserved, the first-level list
// size ceil(rem*alpha) (both products rounded to float before floor/ceil,
// as __float2int_rd/_ru take floats), the second-level cap max(rem, 0).
struct AbmBudget {
    int32_t rem, ins_raw, t1, t2;
};

PG_HD AbmBudget abm_budget(int32_t k, int32_t sy, double alpha, double frc) {
    AbmBudget b;
    const int32_t reserved = (int32_t)__builtin_floorf((float)(frc * (double)k));
    b.rem = k - 1 - sy - reserved;
    b.ins_raw = (int32_t)__builtin_ceilf((float)((double)b.rem * alpha));
    b.t1 = b.ins_raw > 0 ? b.ins_raw : 0;
    b.t2 = b.rem > 0 ? b.rem : 0;
    return b;
}

// Picks and uniform fill of a paper from its list sizes c1 (first level,
// at most t1) and c2 (second level, at most t2): p1, p2, n_uniform.
PG_HD void abm_counts(int32_t k, int32_t sy, const AbmBudget& b, int32_t c1, int32_t c2,
                      int64_t n_exist, int32_t* p1, int32_t* p2, int32_t* nuni) {
    const int32_t inside = b.ins_raw > 0 ? c1 : b.ins_raw;
    const int32_t rem3 = b.rem - inside;
    const int32_t outside = rem3 > 0 ? (rem3 < c2 ? rem3 : c2) : rem3;
    const int32_t fully = k - 1 - sy - inside - outside;
    *p1 = inside > 0 ? inside : 0;
    *p2 = outside > 0 ? outside : 0;
    const int64_t avail = n_exist - 1 - *p1 - *p2;
    *nuni = fully > 0 ? (int32_t)(fully < avail ? fully : avail) : 0;
}

// Selection key of candidate c for citing paper v (src/abm.cu:495-545):
// log(u) / score in FP64, truncated to float, u a float in (0,1]; score 0
// gives the reference's INT_MIN.  Returned as a 64-bit value whose ascending
// order is (key descending, id ascending): the smallest values are kept.
PG_HD uint64_t abm_key(uint64_t seed, uint32_t v, int32_t c, uint32_t yi, float P, float R, float F,
                       const double* w) {
    const u32x4 r = philox(v, (uint32_t)c, yi, TAG_ABM_KEY, seed);
    const float u = (float)((r.x >> 8) + 1u) * 0x1p-24f;
    const float lg = det_logf(u);
    const double s = ((double)P * w[0] + (double)R * w[1]) + (double)F * w[2];
    const float key = s != 0.0 ? (float)((double)lg / s) : -2147483648.0f;
    const uint32_t kb = f2u(key);
    const uint32_t asc = (kb & 0x80000000u) ? ~kb : (kb | 0x80000000u);
    return ((uint64_t)(~asc) << 32) | (uint32_t)c;
}

// PA weight of one paper: exp(alpha*log(k+1) - age*inv_tau + logfit), FP32.
PG_HD float pa_weight(int32_t k, int32_t age, float logfit, float alpha, float inv_tau) {
    float lk = det_logf((float)(k + 1));
    float t1 = __builtin_fmaf(-(float)age, inv_tau, logfit);
    float x = __builtin_fmaf(alpha, lk, t1);
    return det_expf(x);
}

}  // namespace pg
