/*
 * oracle.h -- CPU restatement of ParAGraph's yearly graph-growth step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in paragraph_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to check or to time the
 * CPU restatement.  The product path never links or calls it.
 *
 * It restates, in plain C with OpenMP over the papers of a year, the model
 * that the reference's year loop runs (src/kernel.cu:1682-1845) with the
 * growth model's formulas (FP32 weights (k+1)^a * exp(-age/tau) * fitness, FP64
 * CDF, binary-search draws without duplicates, 2-hop neighbourhood draws,
 * ER G(n,p) by geometric skips, CSR append with degree update).  Every
 * random number is a Philox4x32-10 output keyed by the simulation seed and
 * indexed by (paper, attempt, year index, tag), so the GPU path consumes the
 * identical stream.
 *
 * Parity pinning (what checks each function against something other than
 * itself) is listed in the header comment of oracle.c.
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 53-bit uniform in [0,1) from two words: ((x0<<21) ^ (x1>>11)) * 2^-53. */
double or_u53(uint32_t x0, uint32_t x1);
/* 53-bit uniform in (0,1]. */
double or_u53_open0(uint32_t x0, uint32_t x1);

/* Deterministic elementary functions (same operation sequence on the GPU). */
float or_expf(float x);
float or_logf(float x);
double or_log(double x);
double or_log1p(double x);
void or_vec_expf(int64_t n, const float* x, float* y);
void or_vec_logf(int64_t n, const float* x, float* y);
void or_vec_log(int64_t n, const double* x, double* y);
void or_vec_log1p(int64_t n, const double* x, double* y);

/* Compressed sparse rows of the citation graph: fwd = papers each paper
 * cites, bwd = papers citing each paper; every row sorted ascending. */
typedef struct {
    int64_t n;
    const int32_t* fwd_off;
    const int32_t* fwd_adj;
    const int32_t* bwd_off;
    const int32_t* bwd_adj;
} or_csr;

/* w_i = exp(fma(alpha, log(k_i+1), fma(-(year-birth_i), inv_tau, logfit_i))) */
void or_weights(int64_t n, const int32_t* in_deg, const int32_t* birth,
                const float* logfit, int32_t year, float alpha, float inv_tau,
                float* w_out);
/* Sequential FP64 inclusive prefix sum of the FP32 weights. */
void or_cdf(int64_t n, const float* w, double* cdf_out);
/* First index k with cdf[k] > t, clamped to n-1. */
int64_t or_upper_bound(const double* cdf, int64_t n, double t);

/* 2-hop undirected neighbourhood of root (root excluded, ascending ids,
 * 1-hop frontier and output both capped to the cap smallest ids).
 * Writes at most cap ids to out; returns the count. */
int32_t or_bfs(const or_csr* g, int32_t root, int32_t cap, int32_t* out);

typedef struct {
    uint64_t seed;
    int32_t year_index;   /* 0-based growth year; Philox counter word 2 */
    int32_t year;         /* calendar year given to the new papers */
    int64_t n_exist;      /* papers at the start of the year (ids [0,n_exist)) */
    int64_t n_new;        /* papers added this year (ids n_exist..) */
    int32_t m;            /* citations per new paper */
    int32_t n_nb;         /* of which drawn from the root's 2-hop neighbourhood */
    int32_t cap;          /* BFS frontier / output cap */
    float alpha;
    float inv_tau;
} or_pa_params;

/* One PA year.  Outputs per new paper i: cit[i*m + j] (draw order),
 * kind[i*m + j] (0 PA, 1 neighbourhood, 2 PA refill), cnt[i],
 * near[i] = PA attempts whose uniform fell within 1e-12*Z of a CDF boundary.
 * Returns 0 on success. */
int or_pa_year(const or_csr* g, const int32_t* in_deg, const int32_t* birth,
               const float* logfit, const or_pa_params* p, int32_t* cit,
               int8_t* kind, int32_t* cnt, int32_t* near, int nthreads);

typedef struct {
    uint64_t seed;
    int32_t year_index;
    int64_t n_exist;
    int64_t n_new;
    double p;
} or_er_params;

/* One ER G(n,p) year.  off_out[n_new+1], tgt_out[cap_out].
 * Returns the edge count, or -1 if cap_out is too small. */
int64_t or_er_year(const or_er_params* p, int32_t* off_out, int32_t* tgt_out,
                   int64_t cap_out, int nthreads);

/* CSR append at a year boundary.  New paper i (id n_old+i) cites
 * new_tgt[new_off[i] .. new_off[i+1]) (distinct, any order; existing papers
 * or other new papers of the year).  Output arrays
 * are sized for n_old+n_new rows and E_old+E_added entries; in_deg_out gets
 * the new backward row lengths. */
int or_csr_append(const or_csr* g, int64_t n_new, const int32_t* new_off,
                  const int32_t* new_tgt, int32_t* fwd_off_out,
                  int32_t* fwd_adj_out, int32_t* bwd_off_out,
                  int32_t* bwd_adj_out, int32_t* in_deg_out);

/* FP64 log-likelihood of a PA citation sequence and its gradient.
 * Years y = 0..n_years-1: papers [0, n_exist[y]) exist with in-degree
 * snapshot deg + deg_off[y] (length n_exist[y]); new papers
 * [n_exist[y], n_exist[y]+n_new[y]) carry m slots each in
 * cit/kind + ev_off[y] and counts cnt + cnt_off[y].  Kinds 0 and 2 are
 * likelihood events; kind 1 draws only shrink the partition.
 * grad_phi has one entry per paper (sized n_exist[last]+n_new[last]). */
int or_pa_loglik_grad(int32_t n_years, const int64_t* n_exist,
                      const int64_t* n_new, const int32_t* year_of,
                      const int32_t* deg, const int64_t* deg_off,
                      const int32_t* cit, const int8_t* kind,
                      const int64_t* ev_off, const int32_t* cnt,
                      const int64_t* cnt_off, int32_t m,
                      const int32_t* birth, const float* logfit, double alpha,
                      double tau, double* loglik, double* grad_alpha,
                      double* grad_tau, double* grad_phi);

/* Same sequence and outputs, with the GPU likelihood's arithmetic: L, x and
 * w are FP32 (the simulator's weight formula with alpha and 1/tau rounded to
 * float), partition sums and gradients FP64, log(D) = or_log. */
int or_pa_loglik_grad_w32(int32_t n_years, const int64_t* n_exist,
                          const int64_t* n_new, const int32_t* year_of,
                          const int32_t* deg, const int64_t* deg_off,
                          const int32_t* cit, const int8_t* kind,
                          const int64_t* ev_off, const int32_t* cnt,
                          const int64_t* cnt_off, int32_t m,
                          const int32_t* birth, const float* logfit, double alpha,
                          double tau, double* loglik, double* grad_alpha,
                          double* grad_tau, double* grad_phi);

/* ---- the reference's scoring model (abm_model.c) ---- */
typedef struct {
    uint64_t seed;
    const int32_t* bag;        /* out-degree bag values, each in [0, 250] */
    int32_t n_bag;
    const double* rec_prob;    /* probability of year difference d, d in [0, n_rec) */
    int32_t n_rec;
    double alpha;              /* neighbourhood share in [0,1]; < 0: per-paper uniform */
    double w_pa, w_rec, w_fit; /* score weights; any of them -1: per-paper uniform */
    double frc;                /* fully random share in [0,1] */
    double syp;                /* same-year proportion in [0,1] */
} or_abm_model;

/* Cumulative fitness-peak probabilities of f = 1..1000 (1000 doubles). */
void or_abm_fitness_cdf(double* cdf);
/* Static fitness peak value of paper id, in [1,1000]. */
int32_t or_abm_fitness(uint64_t seed, int64_t id, const double* cdf);
/* Out-degree, neighbourhood share and normalised weights (w[3]) of paper id. */
void or_abm_paper(const or_abm_model* m, int64_t id, int32_t* k, double* alpha, double* w);
/* Score components of papers [0,n) in a year: comp[4c..4c+3] = P, R, F, 0.
 * Returns 0, or 2 if an in-degree reaches 2^21. */
int or_abm_scores(int64_t n, const int32_t* in_deg, const int32_t* birth, int32_t year,
                  const or_abm_model* m, float* comp);
/* Selection key (a float, as double) of candidate c for citing paper v. */
double or_abm_key(const or_abm_model* m, uint32_t v, int32_t c, uint32_t yi, const float* comp,
                  const double* w);
/* Same-year citation target of each of the year's D papers, -1 for none. */
int or_abm_same_year(const or_abm_model* m, int64_t N, int64_t D, uint32_t yi, int32_t* target);
/* One year of the model on graph g (N = g->n papers): new paper l's row is
 * tgt_out[off_out[l] .. off_out[l+1]) ascending; info[6l..6l+5] = generator,
 * out-degree, same-year target (-1), F1 picks, F2 picks, uniform picks;
 * comp_out (optional, [4N]) gets the score components.  Returns the edge
 * count, -1 on failure or if cap_out is too small, -2 if N < 1. */
int64_t or_abm_year(const or_csr* g, const int32_t* in_deg, const int32_t* birth,
                    const or_abm_model* m, int32_t year, int32_t year_index, int64_t n_new,
                    int32_t* off_out, int32_t* tgt_out, int64_t cap_out, float* comp_out,
                    int32_t* info, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
