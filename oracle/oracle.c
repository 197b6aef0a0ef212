/*
 * oracle.c -- CPU restatement of the graph-growth step.  TEST INFRASTRUCTURE
 * ONLY (the checker for the HIP path; never called by the product).
 *
 * The reference's own CPU path (cpu-model branch) is not available, so this
 * file restates the model from the reference's year-loop structure and the
 * growth model's formulas.  Each function names the reference code it
 * follows.
 *
 * Parity pinning -- what checks each part against something independent of
 * this file (tests/test_oracle_*.py):
 *   or_philox4x32_10  Random123 known-answer vectors, and rocRAND's
 *                     philox4x32_10 engine compiled on the host.
 *   or_expf/or_logf/  ulp bounds against correctly rounded values from
 *   or_log/or_log1p   numpy float64/float128 on dense sweeps.
 *   or_weights        closed forms: alpha=0 -> exp(-age/tau)*f, k=0 -> 1*...,
 *                     inv_tau=0 -> (k+1)^alpha * f (to the ulp bounds above).
 *   or_cdf            integer-valued and power-of-two weights (exact sums).
 *   or_pa_year        exact enumeration of without-replacement probabilities
 *                     on tiny graphs (chi-square), invariants (no duplicates,
 *                     no self/new targets, counts), a planted single-weight
 *                     case, and the neighbourhood rule on hand-built graphs.
 *   or_bfs            brute force with Python sets on random graphs plus
 *                     star / path / clique / isolated / hub-over-cap cases.
 *   or_er_year        p=0 and p=1, Binomial(n,p) mean/variance, geometric
 *                     gap law (KS), agreement in law with per-pair Bernoulli.
 *   or_csr_append     transpose and degree-sum invariants against a
 *                     brute-force rebuild from the edge list.
 *   or_pa_loglik_grad closed form at alpha=0, tau=inf, phi=0
 *                     (log L = -sum log(N_y - j)), central finite differences,
 *                     and brute-force enumeration of sequence probabilities.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11.  Replaces the reference's curand XORWOW        */
/* (src/kernel.cu:233, src/abm.cu:529).                                    */
/* ---------------------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
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
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_u53(uint32_t x0, uint32_t x1) {
    uint64_t v = ((uint64_t)x0 << 21) ^ (uint64_t)(x1 >> 11);
    return (double)v * 0x1p-53;
}

double or_u53_open0(uint32_t x0, uint32_t x1) {
    uint64_t v = ((uint64_t)x0 << 21) ^ (uint64_t)(x1 >> 11);
    return (double)(v + 1) * 0x1p-53;
}

static inline void draw4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                         uint32_t c3, uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* ---------------------------------------------------------------------- */
/* Deterministic exp / log.  Cody & Waite argument reduction (Cody, Waite,  */
/* "Software Manual for the Elementary Functions", 1980) with Taylor        */
/* polynomials evaluated by Horner's rule using explicit fma; log uses the  */
/* atanh series log(1+f) = 2 atanh(f/(2+f)).  Every operation is written    */
/* out so the GPU compiles the identical IEEE sequence.                     */
/* ---------------------------------------------------------------------- */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

float or_expf(float x) {
    if (x != x) return x;
    if (x > 88.7228317f) return INFINITY;
    if (x < -87.3365402f) return 0.0f;
    float kf = rintf(x * 0x1.715476p+0f);
    float r = fmaf(kf, -0x1.62e4p-1f, x);
    r = fmaf(kf, -0x1.7f7d1cp-20f, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int n = (int)kf;
    int n1 = n / 2;
    int n2 = n - n1;
    float s1 = u2f((uint32_t)(n1 + 127) << 23);
    float s2 = u2f((uint32_t)(n2 + 127) << 23);
    return (p * s1) * s2;
}

float or_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    uint32_t b = f2u(x);
    int e = 0;
    if (b < 0x00800000u) { x = x * 8388608.0f; b = f2u(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float mm = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (mm > 1.41421354f) { mm = mm * 0.5f; e += 1; }
    float f = mm - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = fmaf(z, 2.0f / 11.0f, 2.0f / 9.0f);
    R = fmaf(z, R, 2.0f / 7.0f);
    R = fmaf(z, R, 2.0f / 5.0f);
    R = fmaf(z, R, 2.0f / 3.0f);
    R = z * R;
    float t = f - R;
    t = s * t;
    float lm = f - t;
    float ef = (float)e;
    return fmaf(ef, 0x1.62e4p-1f, fmaf(ef, 0x1.7f7d1cp-20f, lm));
}

static inline double log_core(double f) {
    double s = f / (2.0 + f);
    double z = s * s;
    double R = fma(z, 2.0 / 21.0, 2.0 / 19.0);
    R = fma(z, R, 2.0 / 17.0);
    R = fma(z, R, 2.0 / 15.0);
    R = fma(z, R, 2.0 / 13.0);
    R = fma(z, R, 2.0 / 11.0);
    R = fma(z, R, 2.0 / 9.0);
    R = fma(z, R, 2.0 / 7.0);
    R = fma(z, R, 2.0 / 5.0);
    R = fma(z, R, 2.0 / 3.0);
    R = z * R;
    double t = f - R;
    t = s * t;
    return f - t;
}

double or_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    uint64_t b = d2u(x);
    int e = 0;
    if (b < 0x0010000000000000ull) { x = x * 0x1p52; b = d2u(x); e = -52; }
    e += (int)(b >> 52) - 1023;
    double mm = u2d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (mm > 1.4142135623730951) { mm = mm * 0.5; e += 1; }
    double lm = log_core(mm - 1.0);
    double ef = (double)e;
    return fma(ef, 6.93147180369123816490e-01, fma(ef, 1.90821492927058770002e-10, lm));
}

double or_log1p(double x) {
    if (x > -0.29289321881345248 && x < 0.41421356237309515) return log_core(x);
    return or_log(1.0 + x);
}

/* ---------------------------------------------------------------------- */
/* Weights and CDF.  Reference score prep: src/abm.cu:112-122 (in-degree),  */
/* src/abm.cu:152-176 (fitness), src/abm.cu:324-338 (scores), frozen at the */
/* start of the year (src/kernel.cu:1698-1706).  Weight (k+1)^a*e^(-age/t)*f. */
/* ---------------------------------------------------------------------- */
static inline float weight_one(int32_t k, int32_t age, float logfit, float alpha,
                               float inv_tau) {
    float lk = or_logf((float)(k + 1));
    float t1 = fmaf(-(float)age, inv_tau, logfit);
    float x = fmaf(alpha, lk, t1);
    return or_expf(x);
}

void or_weights(int64_t n, const int32_t* in_deg, const int32_t* birth,
                const float* logfit, int32_t year, float alpha, float inv_tau,
                float* w_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        w_out[i] = weight_one(in_deg[i], year - birth[i], logfit[i], alpha, inv_tau);
}

void or_cdf(int64_t n, const float* w, double* cdf_out) {
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        acc += (double)w[i];
        cdf_out[i] = acc;
    }
}

int64_t or_upper_bound(const double* cdf, int64_t n, double t) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] > t) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

/* ---------------------------------------------------------------------- */
/* Sorting helpers for the BFS (LSD radix for large lists).                 */
/* ---------------------------------------------------------------------- */
static void isort(int32_t* a, int64_t n) {
    for (int64_t i = 1; i < n; ++i) {
        int32_t v = a[i];
        int64_t j = i - 1;
        while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; --j; }
        a[j + 1] = v;
    }
}

/* Sorts non-negative ints; tmp has room for n. */
static void rsort(int32_t* a, int64_t n, int32_t* tmp) {
    if (n <= 64) { isort(a, n); return; }
    int32_t mx = 0;
    for (int64_t i = 0; i < n; ++i) if (a[i] > mx) mx = a[i];
    int32_t* src = a;
    int32_t* dst = tmp;
    for (int shift = 0; shift < 31 && (mx >> shift) > 0; shift += 8) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[((src[i] >> shift) & 255) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 255]++] = src[i];
        int32_t* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, (size_t)n * sizeof(int32_t));
}

static int64_t uniq(int32_t* a, int64_t n) {
    if (n == 0) return 0;
    int64_t o = 1;
    for (int64_t i = 1; i < n; ++i) if (a[i] != a[o - 1]) a[o++] = a[i];
    return o;
}

static int contains_sorted(const int32_t* a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == x;
}

typedef struct { int32_t* a; int32_t* tmp; int32_t* f1; int64_t cap; int64_t f1cap; } bfs_buf;

static void buf_reserve(bfs_buf* b, int64_t need) {
    if (need <= b->cap) return;
    int64_t c = b->cap ? b->cap : 1024;
    while (c < need) c *= 2;
    b->a = (int32_t*)realloc(b->a, (size_t)c * sizeof(int32_t));
    b->tmp = (int32_t*)realloc(b->tmp, (size_t)c * sizeof(int32_t));
    b->cap = c;
}

static void buf_free(bfs_buf* b) { free(b->a); free(b->tmp); free(b->f1); }

/* 2-hop neighbourhood.  Reference: GetOneAndTwoHopNeighborhood              */
/* (src/abm.cu:577-682) and its warp twin (src/abm.cu:692-883): undirected   */
/* (forward and backward CSR) expansion for two levels from the root, root   */
/* excluded, 1-hop and 2-hop disjoint, ids iterated ascending                */
/* (gpu-std/device_vector.cuh:45-173).  The frontier cap keeps the cap       */
/* smallest ids of the 1-hop set before expanding and of the output.         */
static int32_t bfs_impl(const or_csr* g, int32_t root, int32_t cap, int32_t* out, bfs_buf* b) {
    const int32_t* fo = g->fwd_off; const int32_t* fa = g->fwd_adj;
    const int32_t* bo = g->bwd_off; const int32_t* ba = g->bwd_adj;
    int64_t d1 = (int64_t)(fo[root + 1] - fo[root]) + (bo[root + 1] - bo[root]);
    if (b->f1cap < d1 + 1) {
        b->f1cap = d1 + 1;
        b->f1 = (int32_t*)realloc(b->f1, (size_t)b->f1cap * sizeof(int32_t));
    }
    buf_reserve(b, d1 + 1);
    int32_t* f1 = b->f1;
    int64_t n1 = 0;
    for (int32_t e = fo[root]; e < fo[root + 1]; ++e) f1[n1++] = fa[e];
    for (int32_t e = bo[root]; e < bo[root + 1]; ++e) f1[n1++] = ba[e];
    rsort(f1, n1, b->tmp);
    n1 = uniq(f1, n1);
    {   /* drop root */
        int64_t o = 0;
        for (int64_t i = 0; i < n1; ++i) if (f1[i] != root) f1[o++] = f1[i];
        n1 = o;
    }
    int64_t n1c = n1 < cap ? n1 : cap;
    int64_t tot = 0;
    for (int64_t i = 0; i < n1c; ++i) {
        int32_t v = f1[i];
        tot += (int64_t)(fo[v + 1] - fo[v]) + (bo[v + 1] - bo[v]);
    }
    buf_reserve(b, tot + 1);
    int32_t* c = b->a;
    int64_t nc = 0;
    for (int64_t i = 0; i < n1c; ++i) {
        int32_t v = f1[i];
        for (int32_t e = fo[v]; e < fo[v + 1]; ++e) c[nc++] = fa[e];
        for (int32_t e = bo[v]; e < bo[v + 1]; ++e) c[nc++] = ba[e];
    }
    rsort(c, nc, b->tmp);
    nc = uniq(c, nc);
    /* merge F1c with F2 = cand \ {root} \ F1 */
    int64_t i = 0, j = 0, o = 0;
    while (o < cap && (i < n1c || j < nc)) {
        if (j < nc) {
            int32_t x = c[j];
            if (x == root || contains_sorted(f1, n1, x)) { ++j; continue; }
            if (i < n1c && f1[i] < x) out[o++] = f1[i++];
            else { out[o++] = x; ++j; }
        } else {
            out[o++] = f1[i++];
        }
    }
    return (int32_t)o;
}

int32_t or_bfs(const or_csr* g, int32_t root, int32_t cap, int32_t* out) {
    bfs_buf b = {0};
    int32_t r = bfs_impl(g, root, cap, out, &b);
    buf_free(&b);
    return r;
}

/* ---------------------------------------------------------------------- */
/* One PA year.  Follows the reference's per-paper pipeline: candidate      */
/* draws weighted by the frozen scores (MakePopulateCitations,              */
/* src/abm.cu:461-565, Stage 2/3 src/kernel.cu:170-344), neighbourhood of   */
/* a root paper (Stage 1, src/kernel.cu:17-168; generator src/abm.cu:567),  */
/* shortfall fill without duplicates (src/abm.cu:386-432) and the fixed     */
/* citation budget per paper (src/kernel.cu:230-255).                       */
/* ---------------------------------------------------------------------- */
#define PA_ATTEMPTS_PER_CITE 64
#define NEAR_REL 1e-12

/* Neighbourhood draws use an exact fixed-point CDF: weight w becomes
 * floor(w * 2^(49 - e)) with e = the binary exponent field of the largest
 * weight in the neighbourhood (saturated at 2^50), prefix sums are uint64,
 * and a draw takes the first entry whose prefix exceeds
 * floor(v * Z / 2^53), v the 53-bit Philox integer. */
static inline double nb_scale(uint32_t max_bits) {
    int e = (int)(max_bits >> 23) - 127;
    return u2d((uint64_t)(1023 + 49 - e) << 52);
}

static inline uint64_t nb_fixed(float w, double scale) {
    double q = (double)w * scale;
    return q < 0x1p50 ? (uint64_t)q : (uint64_t)1 << 50;
}

static inline uint64_t nb_target(uint32_t x0, uint32_t x1, uint64_t Z) {
    uint64_t v = ((uint64_t)x0 << 21) ^ (uint64_t)(x1 >> 11);
    return (uint64_t)(((unsigned __int128)v * Z) >> 53);
}

static inline int in_list(const int32_t* a, int n, int32_t x) {
    for (int i = 0; i < n; ++i) if (a[i] == x) return 1;
    return 0;
}

int or_pa_year(const or_csr* g, const int32_t* in_deg, const int32_t* birth,
               const float* logfit, const or_pa_params* p, int32_t* cit,
               int8_t* kind, int32_t* cnt, int32_t* near, int nthreads) {
    const int64_t N = p->n_exist;
    if (N <= 0 || p->m < 0 || p->n_nb < 0 || (p->m > 0 && p->n_nb > p->m - 1) || p->cap <= 0)
        return 1;
    float* w = (float*)malloc((size_t)N * sizeof(float));
    double* C = (double*)malloc((size_t)N * sizeof(double));
    if (!w || !C) { free(w); free(C); return 2; }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    or_weights(N, in_deg, birth, logfit, p->year, p->alpha, p->inv_tau, w);
    or_cdf(N, w, C);
    const double Z = C[N - 1];
    const double eps = NEAR_REL * Z;
    const int m = p->m, n_nb = p->n_nb, n_pa = m - n_nb;
    /* no positive weight (every FP32 weight underflowed to zero): no PA
     * attempts, so the year adds no citations */
    const int a_max = Z > 0.0 ? PA_ATTEMPTS_PER_CITE * m : 0;
    const int b_max = PA_ATTEMPTS_PER_CITE * n_nb;
    const uint32_t yi = (uint32_t)p->year_index;

#pragma omp parallel
    {
        bfs_buf bb = {0};
        int32_t* S = (int32_t*)malloc((size_t)p->cap * sizeof(int32_t));
        uint64_t* Q = (uint64_t*)malloc((size_t)p->cap * sizeof(uint64_t));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < p->n_new; ++i) {
            const uint32_t gid = (uint32_t)(N + i);
            int32_t* row = cit + i * m;
            int8_t* kr = kind + i * m;
            int c = 0, a = 0, nr = 0;
            uint32_t r4[4];
            /* phase 1: PA draws, redraw on duplicates */
            while (c < n_pa && a < a_max) {
                draw4(p->seed, gid, (uint32_t)a, yi, 0u, r4);
                double t = or_u53(r4[0], r4[1]) * Z;
                int64_t k = or_upper_bound(C, N, t);
                if ((k > 0 && t - C[k - 1] <= eps) || (k < N - 1 && C[k] - t <= eps)) nr++;
                ++a;
                if (!in_list(row, c, (int32_t)k)) { row[c] = (int32_t)k; kr[c] = 0; ++c; }
            }
            /* phase 2: neighbourhood of the first pick */
            if (n_nb > 0 && c > 0) {
                int32_t ns = bfs_impl(g, row[0], p->cap, S, &bb);
                int avail = 0;
                for (int32_t s = 0; s < ns; ++s) if (!in_list(row, c, S[s])) ++avail;
                if (avail <= n_nb) {
                    for (int32_t s = 0; s < ns; ++s)
                        if (!in_list(row, c, S[s])) { row[c] = S[s]; kr[c] = 1; ++c; }
                } else {
                    /* neighbourhood CDF in 64-bit fixed point, scaled so the
                     * largest weight of S maps below 2^50: exact and
                     * independent of summation order */
                    uint32_t mb = 0;
                    for (int32_t s = 0; s < ns; ++s) {
                        uint32_t b = f2u(w[S[s]]);
                        if (b > mb) mb = b;
                    }
                    const double scale = nb_scale(mb);
                    uint64_t acc = 0;
                    for (int32_t s = 0; s < ns; ++s) { acc += nb_fixed(w[S[s]], scale); Q[s] = acc; }
                    const uint64_t Zs = acc;
                    int got = 0;
                    for (int bt = 0; got < n_nb && bt < b_max; ++bt) {
                        draw4(p->seed, gid, (uint32_t)bt, yi, 1u, r4);
                        const uint64_t t = nb_target(r4[0], r4[1], Zs);
                        int32_t lo = 0, hi = ns;
                        while (lo < hi) {
                            int32_t mid = (lo + hi) >> 1;
                            if (Q[mid] > t) hi = mid; else lo = mid + 1;
                        }
                        if (lo >= ns) lo = ns - 1;
                        int32_t x = S[lo];
                        if (!in_list(row, c, x)) { row[c] = x; kr[c] = 1; ++c; ++got; }
                    }
                }
            }
            /* phase 3: refill any shortfall with PA draws */
            while (c < m && a < a_max) {
                draw4(p->seed, gid, (uint32_t)a, yi, 0u, r4);
                double t = or_u53(r4[0], r4[1]) * Z;
                int64_t k = or_upper_bound(C, N, t);
                if ((k > 0 && t - C[k - 1] <= eps) || (k < N - 1 && C[k] - t <= eps)) nr++;
                ++a;
                if (!in_list(row, c, (int32_t)k)) { row[c] = (int32_t)k; kr[c] = 2; ++c; }
            }
            for (int j = c; j < m; ++j) { row[j] = -1; kr[j] = -1; }
            cnt[i] = c;
            near[i] = nr;
        }
        free(S); free(Q);
        buf_free(&bb);
    }
    free(w); free(C);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* ER G(n,p) by geometric skips (Batagelj & Brandes, "Efficient generation  */
/* of large random networks", Phys. Rev. E 71, 036113, 2005).  The          */
/* reference's gpu-er branch is not available; its CLI and run logs         */
/* (results/slurm-8640738.out:553-556, results/slurm-8634746.out:73) show   */
/* each new node linking to each existing node with probability p.          */
/* ---------------------------------------------------------------------- */
static int64_t er_paper(const or_er_params* p, double c, int64_t i, int32_t* out, int count_only) {
    const int64_t N = p->n_exist;
    const uint32_t gid = (uint32_t)(N + i);
    int64_t e = 0;
    if (p->p <= 0.0) return 0;
    if (p->p >= 1.0) {
        if (!count_only) for (int64_t t = 0; t < N; ++t) out[t] = (int32_t)t;
        return N;
    }
    double pos = -1.0;
    uint32_t r4[4];
    for (uint32_t j = 0;; ++j) {
        draw4(p->seed, gid, j, (uint32_t)p->year_index, 2u, r4);
        double u = or_u53_open0(r4[0], r4[1]);
        double gap = floor(or_log(u) / c);
        double nxt = pos + 1.0 + gap;
        if (nxt >= (double)N) break;
        pos = nxt;
        if (!count_only) out[e] = (int32_t)pos;
        ++e;
    }
    return e;
}

int64_t or_er_year(const or_er_params* p, int32_t* off_out, int32_t* tgt_out,
                   int64_t cap_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double c = or_log1p(-p->p);
    int64_t* counts = (int64_t*)malloc((size_t)(p->n_new + 1) * sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < p->n_new; ++i) counts[i] = er_paper(p, c, i, NULL, 1);
    int64_t tot = 0;
    for (int64_t i = 0; i < p->n_new; ++i) { off_out[i] = (int32_t)tot; tot += counts[i]; }
    off_out[p->n_new] = (int32_t)tot;
    if (tot > cap_out) { free(counts); return -1; }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < p->n_new; ++i) er_paper(p, c, i, tgt_out + off_out[i], 0);
    free(counts);
    return tot;
}

/* ---------------------------------------------------------------------- */
/* CSR append and degree update.  Reference: edge insertion into the        */
/* forward/backward maps (src/kernel.cu:1766-1811), degree update           */
/* (src/graph.cu:537-618) and the CSR rebuild (src/utils.cu:75-142).        */
/* ---------------------------------------------------------------------- */
int or_csr_append(const or_csr* g, int64_t n_new, const int32_t* new_off,
                  const int32_t* new_tgt, int32_t* fwd_off_out,
                  int32_t* fwd_adj_out, int32_t* bwd_off_out,
                  int32_t* bwd_adj_out, int32_t* in_deg_out) {
    const int64_t n = g->n;
    const int64_t ef = g->fwd_off[n];
    /* forward: old rows verbatim, new rows sorted */
    memcpy(fwd_off_out, g->fwd_off, (size_t)(n + 1) * sizeof(int32_t));
    memcpy(fwd_adj_out, g->fwd_adj, (size_t)ef * sizeof(int32_t));
    int64_t pos = ef;
    for (int64_t i = 0; i < n_new; ++i) {
        int64_t len = new_off[i + 1] - new_off[i];
        memcpy(fwd_adj_out + pos, new_tgt + new_off[i], (size_t)len * sizeof(int32_t));
        isort(fwd_adj_out + pos, len);
        for (int64_t k = 1; k < len; ++k)
            if (fwd_adj_out[pos + k] == fwd_adj_out[pos + k - 1]) return 1;
        pos += len;
        fwd_off_out[n + i + 1] = (int32_t)pos;
    }
    /* backward: counting placement, citers in ascending id order */
    int64_t* add = (int64_t*)calloc((size_t)(n + n_new), sizeof(int64_t));
    for (int64_t i = 0; i < n_new; ++i)
        for (int64_t e = new_off[i]; e < new_off[i + 1]; ++e) {
            /* existing papers, or (same-year citations) other new papers */
            int32_t t = new_tgt[e];
            if (t < 0 || t >= n + n_new || t == n + i) { free(add); return 2; }
            add[t]++;
        }
    bwd_off_out[0] = 0;
    for (int64_t v = 0; v < n + n_new; ++v) {
        int64_t old = v < n ? (int64_t)(g->bwd_off[v + 1] - g->bwd_off[v]) : 0;
        bwd_off_out[v + 1] = (int32_t)(bwd_off_out[v] + old + add[v]);
        in_deg_out[v] = (int32_t)(old + add[v]);
    }
    int64_t* cur = add;
    for (int64_t v = 0; v < n + n_new; ++v) {
        int64_t old = v < n ? (int64_t)(g->bwd_off[v + 1] - g->bwd_off[v]) : 0;
        if (old)
            memcpy(bwd_adj_out + bwd_off_out[v], g->bwd_adj + g->bwd_off[v], (size_t)old * sizeof(int32_t));
        cur[v] = bwd_off_out[v] + old;
    }
    for (int64_t i = 0; i < n_new; ++i)
        for (int64_t e = new_off[i]; e < new_off[i + 1]; ++e)
            bwd_adj_out[cur[new_tgt[e]]++] = (int32_t)(n + i);
    free(add);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* PA log-likelihood and gradient (no reference counterpart).               */
/* log L = sum over PA draws d of [x_t - log(Z_y - sum_{s in B_d} w_s)],    */
/* x_i = alpha*log(k_i+1) - age_i/tau + phi_i, w = exp(x), B_d = papers the */
/* same citing paper cited before draw d.                                   */
/* w32 = 0: every quantity in FP64 (libm exp/log): the smooth function that */
/*          finite differences check.                                       */
/* w32 = 1: L, x and w are the simulator's FP32 values (weight_one's        */
/*          arithmetic with alpha and 1/tau rounded to float); sums,        */
/*          partitions and gradients stay FP64 and log(D) is or_log.  This  */
/*          is the GPU likelihood's arithmetic, up to summation order.      */
/* ---------------------------------------------------------------------- */
static int loglik_impl(int32_t n_years, const int64_t* n_exist,
                       const int64_t* n_new, const int32_t* year_of,
                       const int32_t* deg, const int64_t* deg_off,
                       const int32_t* cit, const int8_t* kind,
                       const int64_t* ev_off, const int32_t* cnt,
                       const int64_t* cnt_off, int32_t m,
                       const int32_t* birth, const float* logfit, double alpha,
                       double tau, int w32, double* loglik, double* grad_alpha,
                       double* grad_tau, double* grad_phi) {
    int64_t nmax = 0;
    for (int y = 0; y < n_years; ++y)
        if (n_exist[y] + n_new[y] > nmax) nmax = n_exist[y] + n_new[y];
    double* w = (double*)malloc((size_t)(nmax > 0 ? nmax : 1) * sizeof(double));
    double* L = (double*)malloc((size_t)(nmax > 0 ? nmax : 1) * sizeof(double));
    double* X = (double*)malloc((size_t)(nmax > 0 ? nmax : 1) * sizeof(double));
    double* later = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    memset(grad_phi, 0, (size_t)nmax * sizeof(double));
    const float alpha_f = (float)alpha;
    const float inv_tau_f = (float)(1.0 / tau);
    double ll = 0.0, ga = 0.0, gt = 0.0;
    const double tau2 = tau * tau;
    for (int y = 0; y < n_years; ++y) {
        const int64_t N = n_exist[y];
        const int32_t* k = deg + deg_off[y];
        double Z = 0.0, SL = 0.0, Sa = 0.0;
        for (int64_t j = 0; j < N; ++j) {
            const int32_t age_i = year_of[y] - birth[j];
            const double age = (double)age_i;
            if (w32) {
                const float lf = or_logf((float)(k[j] + 1));
                const float xf = fmaf(alpha_f, lf, fmaf(-(float)age_i, inv_tau_f, logfit[j]));
                L[j] = (double)lf;
                X[j] = (double)xf;
                w[j] = (double)or_expf(xf);
            } else {
                L[j] = log((double)k[j] + 1.0);
                X[j] = alpha * L[j] - age / tau + (double)logfit[j];
                w[j] = exp(X[j]);
            }
            Z += w[j]; SL += w[j] * L[j]; Sa += w[j] * age;
        }
        double A = 0.0;
        for (int64_t i = 0; i < n_new[y]; ++i) {
            const int32_t* row = cit + ev_off[y] + i * m;
            const int8_t* kr = kind + ev_off[y] + i * m;
            const int32_t c = cnt[cnt_off[y] + i];
            double Pw = 0.0, PwL = 0.0, Pwa = 0.0;
            for (int j = 0; j < c; ++j) later[j] = 0.0;
            for (int j = 0; j < c; ++j) {
                const int32_t t = row[j];
                const double age_t = (double)(year_of[y] - birth[t]);
                if (kr[j] != 1) {
                    double D = Z - Pw;
                    ll += X[t] - (w32 ? or_log(D) : log(D));
                    A += 1.0 / D;
                    ga += L[t] + PwL / D;
                    gt += age_t / tau2 + Pwa / (tau2 * D);
                    grad_phi[t] += 1.0;
                    for (int q = 0; q < j; ++q) later[q] += 1.0 / D;
                }
                Pw += w[t]; PwL += w[t] * L[t]; Pwa += w[t] * age_t;
            }
            for (int q = 0; q < c; ++q) grad_phi[row[q]] += w[row[q]] * later[q];
        }
        ga -= A * SL;
        gt -= A * Sa / tau2;
        for (int64_t j = 0; j < N; ++j) grad_phi[j] -= A * w[j];
    }
    *loglik = ll; *grad_alpha = ga; *grad_tau = gt;
    free(w); free(L); free(X); free(later);
    return 0;
}

int or_pa_loglik_grad(int32_t n_years, const int64_t* n_exist,
                      const int64_t* n_new, const int32_t* year_of,
                      const int32_t* deg, const int64_t* deg_off,
                      const int32_t* cit, const int8_t* kind,
                      const int64_t* ev_off, const int32_t* cnt,
                      const int64_t* cnt_off, int32_t m,
                      const int32_t* birth, const float* logfit, double alpha,
                      double tau, double* loglik, double* grad_alpha,
                      double* grad_tau, double* grad_phi) {
    return loglik_impl(n_years, n_exist, n_new, year_of, deg, deg_off, cit, kind, ev_off, cnt,
                       cnt_off, m, birth, logfit, alpha, tau, 0, loglik, grad_alpha, grad_tau,
                       grad_phi);
}

int or_pa_loglik_grad_w32(int32_t n_years, const int64_t* n_exist,
                          const int64_t* n_new, const int32_t* year_of,
                          const int32_t* deg, const int64_t* deg_off,
                          const int32_t* cit, const int8_t* kind,
                          const int64_t* ev_off, const int32_t* cnt,
                          const int64_t* cnt_off, int32_t m,
                          const int32_t* birth, const float* logfit, double alpha,
                          double tau, double* loglik, double* grad_alpha,
                          double* grad_tau, double* grad_phi) {
    return loglik_impl(n_years, n_exist, n_new, year_of, deg, deg_off, cit, kind, ev_off, cnt,
                       cnt_off, m, birth, logfit, alpha, tau, 1, loglik, grad_alpha, grad_tau,
                       grad_phi);
}

void or_vec_expf(int64_t n, const float* x, float* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = or_expf(x[i]);
}
void or_vec_logf(int64_t n, const float* x, float* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = or_logf(x[i]);
}
void or_vec_log(int64_t n, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = or_log(x[i]);
}
void or_vec_log1p(int64_t n, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = or_log1p(x[i]);
}
