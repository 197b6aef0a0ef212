/*
 * abm_model.c -- CPU restatement of the reference's scoring model (the
 * agent-based model of src/abm.cu driven by src/kernel.cu:1682-1845).  TEST
 * INFRASTRUCTURE ONLY: the checker for the HIP path's ABM year; the product
 * never links or calls it.
 *
 * Per year with N existing papers and D new ones (ids N..N+D-1):
 *   scores of every existing paper c (frozen for the year,
 *   src/kernel.cu:1698-1706):
 *     P_c = (k_c^3 + 1) / sum            in-degree k_c, gamma = 3
 *                                        (CalculateScores, src/abm.cu:324-338)
 *     F_c = (f_c^3 + 1) / sum            static fitness peak value f_c in
 *                                        [1,1000], P(f) ~ f^-1.634
 *                                        (src/abm.cu:124-150, :152-176;
 *                                        includes/abm.cuh:469-484)
 *     R_c = p(d_c) / (count(d_c) * U)    d_c = year - birth_c, p from the
 *                                        recency table, U = sum of p over the
 *                                        differences present
 *                                        (FillRecencyArr, src/abm.cu:179-208)
 *   each new paper v: out-degree k from the bag (src/abm.cu:282-290, capped
 *   at 250, the reference's citation buffer, src/kernel.cu:1151-1153), share
 *   alpha and weights (src/abm.cu:241-280), a uniform generator g among the
 *   existing papers (src/abm.cu:567-574); a fixed-size uniform subset of the
 *   year's papers (floor(D * same_year_proportion) of them,
 *   src/abm.cu:340-353) cites one other new paper (src/abm.cu:376-384);
 *   budgets (src/kernel.cu:244-255, :383-386):
 *     reserved = floor(frc * k), rem = k - 1 - sy - reserved,
 *     inside = min(ceil(rem * alpha), |F1|), rem3 = rem - inside,
 *     outside = min(rem3, |F2|), fully = k - 1 - sy - inside - outside
 *   with F1 = (fwd(g) U bwd(g)) \ {g}, F2 = (U_{u in F1} fwd(u) U bwd(u)) \
 *   F1 \ {g} (src/abm.cu:577-690); `inside` papers of F1 and `outside` of F2
 *   by Efraimidis-Spirakis keys log(u) / score (src/abm.cu:461-565), score =
 *   P_c*w_pa + R_c*w_rec + F_c*w_fit with the components read as float
 *   (src/abm.cu:499-501), the float key truncation of src/abm.cu:540; then
 *   `fully` distinct uniform papers other than g and the picks
 *   (src/abm.cu:386-432); the edge to g always (src/abm.cu:990).
 *
 * Randomness: Philox4x32-10 under the simulation seed, counter (id, word,
 * year word, tag); this file keeps its own copy of every draw's definition
 * (the product's lives in paragraph_amd/csrc/pg_math.h).  Where the
 * reference has a quirk that only a bug explains, this model does not
 * follow it: the same-year citation never targets the citing paper itself,
 * the uniform fill can pick every existing paper (the reference's
 * `curand % (N-1)` never picks N-1 and has modulo bias), and the fill's
 * availability bound counts existing papers only.
 *
 * Parity pinning (tests/test_oracle_abm.py): the score components against
 * closed forms (sums of P, F and R are 1; hand-computed small cases); the
 * fitness law against f^-1.634 (chi-square); per-paper out-degree against
 * the bag (chi-square), alpha and weights (uniform / normalisation);
 * F1 and F2 against brute-force Python sets; the key selection against the
 * exact successive-sampling law of weighted sampling without replacement on
 * tiny candidate sets (chi-square over ordered pairs); the budget formulas
 * against hand-derived edge counts of the reference's arithmetic for the
 * boundary cases (k = 0, 1; same-year; empty neighbourhoods; alpha 0 and 1);
 * the same-year subset size and uniformity; the uniform fill's exclusions.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { T_FIT = 7, T_PAPER = 8, T_GEN = 9, T_SAME = 10, T_KEY = 11, T_UNI = 12 };
#define PAPER_WORD 0xFFFFFFFFu
#define KMAX 250

static void ph(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t o[4]) {
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, o);
}

/* floor(u * n) for a 53-bit u in [0,1), clamped to n-1. */
static int64_t pick(uint32_t a, uint32_t b, int64_t n) {
    int64_t x = (int64_t)floor(or_u53(a, b) * (double)n);
    return x >= n ? n - 1 : x;
}

void or_abm_fitness_cdf(double* cdf) {
    double s = 0.0;
    for (int f = 1; f <= 1000; ++f) {
        s += 6.3742991333 * 0.072 * pow((double)f, -1.634);
        cdf[f - 1] = s;
    }
}

int32_t or_abm_fitness(uint64_t seed, int64_t id, const double* cdf) {
    uint32_t r[4];
    ph(seed, (uint32_t)id, 0u, PAPER_WORD, T_FIT, r);
    const double t = or_u53(r[0], r[1]) * cdf[999];
    int lo = 0, hi = 999;  /* first index with cdf > t */
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1;
    }
    return lo + 1;
}

void or_abm_paper(const or_abm_model* m, int64_t id, int32_t* k, double* alpha, double* w) {
    uint32_t r[4];
    ph(m->seed, (uint32_t)id, 0u, PAPER_WORD, T_PAPER, r);
    *k = m->bag[pick(r[0], r[1], m->n_bag)];
    *alpha = m->alpha >= 0.0 ? m->alpha : or_u53(r[2], r[3]);
    double a, b, c;
    if (m->w_pa != -1.0 && m->w_rec != -1.0 && m->w_fit != -1.0) {
        a = m->w_pa; b = m->w_rec; c = m->w_fit;
    } else {
        uint32_t q[4];
        ph(m->seed, (uint32_t)id, 1u, PAPER_WORD, T_PAPER, r);
        ph(m->seed, (uint32_t)id, 2u, PAPER_WORD, T_PAPER, q);
        a = or_u53(r[0], r[1]); b = or_u53(r[2], r[3]); c = or_u53(q[0], q[1]);
    }
    const double s = a + b + c;
    if (s > 0.0) {
        w[0] = a / s; w[1] = b / s; w[2] = c / s;
    } else {
        w[0] = w[1] = w[2] = 1.0 / 3.0;
    }
}

int or_abm_scores(int64_t n, const int32_t* in_deg, const int32_t* birth, int32_t year,
                  const or_abm_model* m, float* comp) {
    double cdf[1000];
    or_abm_fitness_cdf(cdf);
    unsigned __int128 sp = 0;
    uint64_t sf = 0;
    int64_t* cnt = (int64_t*)calloc((size_t)(m->n_rec > 0 ? m->n_rec : 1), sizeof(int64_t));
    int32_t* fit = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    if (!cnt || !fit) { free(cnt); free(fit); return 1; }
    for (int64_t c = 0; c < n; ++c) {
        const uint64_t k = (uint64_t)in_deg[c];
        if (k >= (1u << 21)) { free(cnt); free(fit); return 2; }
        sp += (unsigned __int128)(k * k * k + 1);
        fit[c] = or_abm_fitness(m->seed, c, cdf);
        const uint64_t f = (uint64_t)fit[c];
        sf += f * f * f + 1;
        const int64_t d = (int64_t)year - birth[c];
        if (d >= 0 && d < m->n_rec) cnt[d] += 1;
    }
    const double spd = (double)(uint64_t)(sp >> 64) * 0x1p64 + (double)(uint64_t)sp;
    const double sfd = (double)sf;
    double U = 0.0;
    for (int32_t d = 0; d < m->n_rec; ++d)
        if (cnt[d] > 0) U += m->rec_prob[d];
    for (int64_t c = 0; c < n; ++c) {
        const uint64_t k = (uint64_t)in_deg[c];
        const uint64_t f = (uint64_t)fit[c];
        const int64_t d = (int64_t)year - birth[c];
        float r = 0.0f;
        if (d >= 0 && d < m->n_rec && U > 0.0)
            r = (float)(m->rec_prob[d] / ((double)cnt[d] * U));
        comp[4 * c + 0] = (float)((double)(k * k * k + 1) / spd);
        comp[4 * c + 1] = r;
        comp[4 * c + 2] = (float)((double)(f * f * f + 1) / sfd);
        comp[4 * c + 3] = 0.0f;
    }
    free(cnt);
    free(fit);
    return 0;
}

/* Efraimidis-Spirakis key of candidate c for citing paper v, as a 64-bit
 * value whose ascending order is (key descending, id ascending). */
static uint64_t key_of(const or_abm_model* m, uint32_t v, int32_t c, uint32_t yi,
                       const float* comp, const double* w) {
    uint32_t r[4];
    ph(m->seed, v, (uint32_t)c, yi, T_KEY, r);
    const float u = (float)((r[0] >> 8) + 1u) * 0x1p-24f;
    const float lg = or_logf(u);
    const double s = ((double)comp[4 * c] * w[0] + (double)comp[4 * c + 1] * w[1]) +
                     (double)comp[4 * c + 2] * w[2];
    const float key = s != 0.0 ? (float)((double)lg / s) : -2147483648.0f;
    uint32_t b;
    memcpy(&b, &key, 4);
    const uint32_t asc = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((uint64_t)(~asc) << 32) | (uint32_t)c;
}

double or_abm_key(const or_abm_model* m, uint32_t v, int32_t c, uint32_t yi, const float* comp,
                  const double* w) {
    const uint64_t kv = key_of(m, v, c, yi, comp, w);
    const uint32_t asc = ~(uint32_t)(kv >> 32);
    const uint32_t b = (asc & 0x80000000u) ? (asc & 0x7fffffffu) : ~asc;
    float key;
    memcpy(&key, &b, 4);
    return (double)key;
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}
static int cmp_i32(const void* a, const void* b) {
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return x < y ? -1 : x > y;
}

static int in_sorted(const int32_t* a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == x;
}

typedef struct {
    int32_t* f1;   /* distinct F1, ascending */
    uint64_t* kv;  /* keys of the level being selected */
    int32_t* mark; /* per-paper stamp over [0, N): F2 dedup */
    int64_t cap_f1, cap_kv;
} abm_buf;

static int grow32(int32_t** a, int64_t* cap, int64_t need) {
    if (need <= *cap) return 0;
    int64_t c = need > 2 * *cap ? need : 2 * *cap;
    int32_t* p = (int32_t*)realloc(*a, (size_t)c * sizeof(int32_t));
    if (!p) return 1;
    *a = p; *cap = c;
    return 0;
}
static int grow64(uint64_t** a, int64_t* cap, int64_t need) {
    if (need <= *cap) return 0;
    int64_t c = need > 2 * *cap ? need : 2 * *cap;
    uint64_t* p = (uint64_t*)realloc(*a, (size_t)c * sizeof(uint64_t));
    if (!p) return 1;
    *a = p; *cap = c;
    return 0;
}

/* The t smallest distinct 64-bit keys of kv[0,n) into out (their ids),
 * returns their number. */
static int32_t top_ids(uint64_t* kv, int64_t n, int32_t t, int32_t* out) {
    if (t <= 0 || n == 0) return 0;
    qsort(kv, (size_t)n, sizeof(uint64_t), cmp_u64);
    int32_t c = 0;
    for (int64_t i = 0; i < n && c < t; ++i)
        if (i == 0 || kv[i] != kv[i - 1]) out[c++] = (int32_t)(uint32_t)kv[i];
    return c;
}

/* One paper; row gets at most KMAX+1 ids, sorted ascending.  Returns the
 * row length or -1 on allocation failure. */
static int32_t abm_paper(const or_csr* g, const or_abm_model* m, const float* comp,
                         int64_t N, int64_t l, int32_t sy_target, uint32_t yi,
                         abm_buf* b, int32_t* row, int32_t* info) {
    const uint32_t v = (uint32_t)(N + l);
    int32_t k;
    double alpha, w[3];
    or_abm_paper(m, v, &k, &alpha, w);
    uint32_t r[4];
    ph(m->seed, v, 0u, yi, T_GEN, r);
    const int32_t gen = (int32_t)pick(r[0], r[1], N);
    const int32_t sy = sy_target >= 0 ? 1 : 0;
    const int32_t reserved = (int32_t)floorf((float)(m->frc * (double)k));
    const int32_t rem = k - 1 - sy - reserved;
    const int32_t ins_raw = (int32_t)ceilf((float)((double)rem * alpha));
    const int32_t t1 = ins_raw > 0 ? ins_raw : 0;
    const int32_t t2 = rem > 0 ? rem : 0;
    /* F1: both rows of the generator, merged, without the generator */
    const int32_t fa = g->fwd_off[gen], fb = g->fwd_off[gen + 1];
    const int32_t ba = g->bwd_off[gen], bb = g->bwd_off[gen + 1];
    if (grow32(&b->f1, &b->cap_f1, (int64_t)(fb - fa) + (bb - ba) + 1)) return -1;
    int64_t nf1 = 0;
    for (int32_t e = fa; e < fb; ++e) b->f1[nf1++] = g->fwd_adj[e];
    for (int32_t e = ba; e < bb; ++e) b->f1[nf1++] = g->bwd_adj[e];
    qsort(b->f1, (size_t)nf1, sizeof(int32_t), cmp_i32);
    int64_t u1 = 0;
    for (int64_t i = 0; i < nf1; ++i)
        if ((i == 0 || b->f1[i] != b->f1[i - 1]) && b->f1[i] != gen) b->f1[u1++] = b->f1[i];
    nf1 = u1;
    int32_t sel1[KMAX], sel2[KMAX];
    int32_t c1 = 0, c2 = 0;
    if (t1 > 0) {
        if (grow64(&b->kv, &b->cap_kv, nf1 + 1)) return -1;
        for (int64_t i = 0; i < nf1; ++i) b->kv[i] = key_of(m, v, b->f1[i], yi, comp, w);
        c1 = top_ids(b->kv, nf1, t1, sel1);
    }
    if (t2 > 0) {
        /* F2 with a stamp per paper: each candidate keyed once */
        const int32_t stamp = (int32_t)l + 1;
        int64_t n2 = 0;
        for (int64_t i = 0; i < nf1; ++i) b->mark[b->f1[i]] = stamp;
        b->mark[gen] = stamp;
        for (int64_t i = 0; i < nf1; ++i) {
            const int32_t u = b->f1[i];
            for (int pass = 0; pass < 2; ++pass) {
                const int32_t* off = pass ? g->bwd_off : g->fwd_off;
                const int32_t* adj = pass ? g->bwd_adj : g->fwd_adj;
                for (int32_t e = off[u]; e < off[u + 1]; ++e) {
                    const int32_t c = adj[e];
                    if (b->mark[c] == stamp) continue;
                    b->mark[c] = stamp;
                    if (grow64(&b->kv, &b->cap_kv, n2 + 1)) return -1;
                    b->kv[n2++] = key_of(m, v, c, yi, comp, w);
                }
            }
        }
        c2 = top_ids(b->kv, n2, t2, sel2);
        /* the stamps stay; F1 members were marked too, which is harmless
         * since every later paper uses its own stamp */
    }
    const int32_t inside = ins_raw > 0 ? c1 : ins_raw;
    const int32_t rem3 = rem - inside;
    const int32_t outside = rem3 > 0 ? (rem3 < c2 ? rem3 : c2) : rem3;
    const int32_t fully = k - 1 - sy - inside - outside;
    const int32_t p1 = inside > 0 ? inside : 0, p2 = outside > 0 ? outside : 0;
    const int64_t avail = N - 1 - p1 - p2;
    const int32_t nuni = fully > 0 ? (int32_t)(fully < avail ? fully : avail) : 0;
    /* the set so far (existing papers only), ascending, for the fill */
    int32_t n = 0;
    row[n++] = gen;
    for (int32_t i = 0; i < p1; ++i) row[n++] = sel1[i];
    for (int32_t i = 0; i < p2; ++i) row[n++] = sel2[i];
    qsort(row, (size_t)n, sizeof(int32_t), cmp_i32);
    for (uint32_t a = 0, got = 0; got < (uint32_t)nuni; ++a) {
        ph(m->seed, v, a, yi, T_UNI, r);
        const int32_t c = (int32_t)pick(r[0], r[1], N);
        if (in_sorted(row, n, c)) continue;
        int32_t j = n;  /* insertion keeps the row sorted */
        while (j > 0 && row[j - 1] > c) { row[j] = row[j - 1]; --j; }
        row[j] = c;
        ++n;
        ++got;
    }
    if (sy) row[n++] = sy_target;  /* a new paper: past every existing id */
    info[0] = gen; info[1] = k; info[2] = sy_target;
    info[3] = p1; info[4] = p2; info[5] = nuni;
    return n;
}

typedef struct { uint32_t key; int32_t l; } sy_key;
static int cmp_sy(const void* a, const void* b) {
    const sy_key* x = (const sy_key*)a;
    const sy_key* y = (const sy_key*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->l < y->l ? -1 : x->l > y->l;
}

int or_abm_same_year(const or_abm_model* m, int64_t N, int64_t D, uint32_t yi, int32_t* target) {
    for (int64_t l = 0; l < D; ++l) target[l] = -1;
    const int64_t K = (int64_t)floor((double)D * m->syp);
    if (D < 2 || K <= 0) return 0;
    sy_key* s = (sy_key*)malloc((size_t)D * sizeof(sy_key));
    if (!s) return 1;
    uint32_t r[4];
    for (int64_t l = 0; l < D; ++l) {
        ph(m->seed, (uint32_t)(N + l), 0u, yi, T_SAME, r);
        s[l].key = r[0];
        s[l].l = (int32_t)l;
    }
    qsort(s, (size_t)D, sizeof(sy_key), cmp_sy);
    for (int64_t q = 0; q < K && q < D; ++q) {
        const int64_t l = s[q].l;
        ph(m->seed, (uint32_t)(N + l), 0u, yi, T_SAME, r);
        const int64_t j = pick(r[2], r[3], D - 1);
        target[l] = (int32_t)(N + (j < l ? j : j + 1));
    }
    free(s);
    return 0;
}

int64_t or_abm_year(const or_csr* g, const int32_t* in_deg, const int32_t* birth,
                    const or_abm_model* m, int32_t year, int32_t year_index, int64_t n_new,
                    int32_t* off_out, int32_t* tgt_out, int64_t cap_out, float* comp_out,
                    int32_t* info, int nthreads) {
    const int64_t N = g->n, D = n_new;
    if (N < 1) return -2;
    float* comp = comp_out ? comp_out : (float*)malloc((size_t)N * 4 * sizeof(float));
    int32_t* syt = (int32_t*)malloc((size_t)(D > 0 ? D : 1) * sizeof(int32_t));
    int32_t* rows = (int32_t*)malloc((size_t)(D > 0 ? D : 1) * (KMAX + 2) * sizeof(int32_t));
    int32_t* len = (int32_t*)malloc((size_t)(D > 0 ? D : 1) * sizeof(int32_t));
    int fail = !comp || !syt || !rows || !len;
    if (!fail) fail = or_abm_scores(N, in_deg, birth, year, m, comp) ? 1 : 0;
    if (!fail) fail = or_abm_same_year(m, N, D, (uint32_t)year_index, syt);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        abm_buf b = {NULL, NULL, (int32_t*)malloc((size_t)N * sizeof(int32_t)), 0, 0};
        if (b.mark) memset(b.mark, 0, (size_t)N * sizeof(int32_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t l = 0; l < D; ++l) {
            if (fail || !b.mark) { fail = 1; continue; }
            const int32_t r = abm_paper(g, m, comp, N, l, syt[l], (uint32_t)year_index, &b,
                                        rows + l * (KMAX + 2), info + 6 * l);
            if (r < 0) fail = 1; else len[l] = r;
        }
        free(b.f1); free(b.kv); free(b.mark);
    }
    int64_t tot = -1;
    if (!fail) {
        tot = 0;
        off_out[0] = 0;
        for (int64_t l = 0; l < D; ++l) {
            if (tot + len[l] <= cap_out)
                memcpy(tgt_out + tot, rows + l * (KMAX + 2), (size_t)len[l] * sizeof(int32_t));
            tot += len[l];
            off_out[l + 1] = (int32_t)tot;
        }
        if (tot > cap_out) tot = -1;
    }
    if (!comp_out) free(comp);
    free(syt); free(rows); free(len);
    return tot;
}
