/*
 * paragraph.h -- C ABI of the MI355X graph-growth library (libparagraph.so).
 *
 * The reference (ParAGraph, one CUDA executable) has no library or FFI
 * layer.  Its hot path sits behind two boundaries, and this ABI replaces the
 * inner one:
 *   - the per-year host->GPU call buildOneNodeConnections_timed(ABM*, Graph*,
 *     new_nodes, ..., year, N, N0, N_final, max_batch, EpochTiming*)
 *     (src/kernel.cu:1125-1139, declared includes/int2.cuh), which uploads
 *     the score arrays, rebuilds the CSR, runs Stages 1-4 and downloads the
 *     new edges every year;
 *   - the host-side year boundary that inserts those edges and updates the
 *     degrees (src/kernel.cu:1766-1815, src/graph.cu:537-618).
 * The outer boundary (the `abm` command line and its edgelist/nodelist files,
 * src/main.cu:29-114, src/graph.cu:336-447, src/graph.cu:909-1218) is kept by
 * the `abm` executable built from paragraph_amd/csrc/abm_main.cpp on top of
 * this ABI.
 *
 * Conventions: every function returns 0 on success or a PG_E* code, and
 * pg_last_error() describes the failure; no C++ exception crosses the ABI.
 * Structs start with a `size` field (set it to sizeof the struct).  The
 * context owns all device memory; callers own host buffers.  A context is
 * used by one host thread at a time.
 */
#ifndef PARAGRAPH_H
#define PARAGRAPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_ABI_VERSION 1

enum {
    PG_OK = 0,
    PG_EINVAL = 1,   /* bad argument or struct size */
    PG_ENOMEM = 2,   /* device or host allocation failed */
    PG_EHIP = 3,     /* HIP runtime error */
    PG_ECAP = 4,     /* a capacity fixed at pg_ctx_create would be exceeded */
    PG_ESTATE = 5,   /* call out of order (e.g. no graph loaded) */
    PG_ENODEV = 6,   /* no usable gfx950 device */
    PG_EIO = 7       /* checkpoint file unreadable, unwritable or malformed */
};

enum {
    PG_MODEL_PA = 0,   /* preferential attachment with ageing and fitness */
    PG_MODEL_ER = 1,   /* Erdos-Renyi G(n,p) */
    PG_MODEL_ABM = 2   /* the reference's scoring model (pg_abm_configure) */
};

typedef struct pg_ctx pg_ctx;

typedef struct {
    uint32_t size;
    int32_t device;            /* HIP device ordinal */
    int64_t max_papers;        /* final graph size the context can hold */
    int64_t max_edges;         /* citation edges the context can hold */
    int64_t max_new_per_year;  /* papers one year may add */
    int32_t max_cites;         /* citations per paper, upper bound (<= 64) */
    int32_t bfs_cap;           /* 2-hop frontier/output cap (<= 4096) */
    uint64_t seed;             /* Philox key of this simulation */
} pg_config;

typedef struct {
    uint32_t size;
    int32_t model;             /* PG_MODEL_PA, PG_MODEL_ER or PG_MODEL_ABM */
    int32_t year;              /* calendar year given to the new papers */
    int32_t year_index;        /* 0-based growth year (Philox counter word) */
    int64_t n_new;             /* papers added this year */
    int32_t m;                 /* PA: citations per new paper */
    int32_t n_nb;              /* PA: of which neighbourhood draws (<= m-1) */
    float alpha;               /* PA: in-degree exponent */
    float inv_tau;             /* PA: 1/tau in 1/years (0 disables ageing) */
    double er_p;               /* ER: link probability */
} pg_year_params;

typedef struct {
    uint32_t size;
    int64_t n_before;          /* papers at the start of the year */
    int64_t n_after;
    int64_t edges_added;
    int64_t near_boundary;     /* PA attempts within 1e-12*Z of a CDF boundary */
    int64_t shortfall;         /* citations not placed (attempt budget spent) */
    int64_t pa_attempts;       /* PA draw attempts consumed (phase 1 + refill) */
    double device_ms;          /* GPU time of the step (hipEvents) */
} pg_year_stats;

/* Kernel groups reported by pg_profile_read. */
enum {
    PG_K_WEIGHTS_CDF = 0,      /* fused weight pass + FP64 CDF */
    PG_K_PA_DRAW = 1,          /* PA draws without duplicates */
    PG_K_BFS = 2,              /* 2-hop BFS of the distinct roots (both paths) */
    PG_K_NB_DRAW = 3,          /* root dedup + neighbourhood draws + refill */
    PG_K_ER = 4,               /* ER count + write */
    PG_K_CSR_APPEND = 5,       /* forward append, radix group, backward merge */
    PG_K_LIK = 6,              /* pg_pa_loglik evaluations (all passes) */
    PG_K_LIK_GRAD = 7,         /* pg_pa_loglik_grad evaluations (all passes) */
    PG_K_LIK_BATCH = 8,        /* pg_pa_loglik_grad_batch calls (all passes) */
    PG_K_ABM_SCORES = 9,       /* scoring model: fitness, score sums and components, same-year subset */
    PG_K_ABM_SELECT = 10,      /* scoring model: 1-hop and 2-hop weighted selection */
    PG_K_ABM_ROWS = 11,        /* scoring model: row counts, scan, uniform fill, row writes */
    PG_K_COUNT = 12
};

/* Returns PG_ABI_VERSION. */
int pg_version(void);

/* Creates a context on cfg->device and allocates its device arena. */
int pg_ctx_create(const pg_config* cfg, pg_ctx** out);
void pg_ctx_destroy(pg_ctx* ctx);
/* Message of the last failure on ctx (or of the last pg_ctx_create failure
 * when ctx is NULL).  Valid until the next call on ctx. */
const char* pg_last_error(const pg_ctx* ctx);

/*
 * Loads the seed graph: n papers with birth years, log-fitness for
 * n_logfit >= n papers (entries beyond n are used by papers added later),
 * and n_edges citations src[i] -> dst[i] (citing -> cited, ids in [0,n)).
 * Duplicate edges are dropped and self-citations rejected, as the
 * reference's set-based adjacency does (src/graph.cu:820-826).
 * Replaces Graph::ParseEdgelist's AddEdge loop plus prepareGraph's CSR build
 * (src/graph.cu:376-447, src/utils.cu:75-142).
 */
int pg_graph_load(pg_ctx* ctx, int64_t n, const int32_t* birth,
                  const float* logfit, int64_t n_logfit, int64_t n_edges,
                  const int32_t* src, const int32_t* dst);

/*
 * Runs one simulated year on the device: PA (weights, FP64 CDF, draws,
 * 2-hop neighbourhood draws, refill) or ER G(n,p), then the CSR append and
 * degree update.  Replaces buildOneNodeConnections_timed
 * (src/kernel.cu:1125-1586) plus the year-boundary edge insertion and
 * degree update (src/kernel.cu:1766-1815).  stats may be NULL.
 */
int pg_year_step(pg_ctx* ctx, const pg_year_params* params, pg_year_stats* stats);

/* ---- the reference's scoring model ----
 *
 * The agent-based model of the reference (src/abm.cu, year loop
 * src/kernel.cu:1682-1845): each new paper gets an out-degree k from the
 * bag, a neighbourhood share alpha and three score weights; it cites a
 * uniformly drawn existing paper (the generator), one other paper of the
 * same year if it is in the year's same-year subset (floor(D * same_year)
 * papers), then ceil(rem * alpha) papers of the generator's 1-hop
 * neighbourhood and the rest of rem = k - 1 - same-year - floor(k *
 * fully_random) from its 2-hop neighbourhood, both by weighted sampling
 * without replacement on score = w_pa * P + w_rec * R + w_fit * F (P: share
 * of in-degree^3 + 1, R: recency table share of the paper's age, F: share
 * of fitness^3 + 1, fitness a static power-law peak value), and fills the
 * remainder with distinct uniform papers.  Replaces the reference's score
 * preparation (src/abm.cu:112-208, :324-338), per-paper draws
 * (src/abm.cu:241-290), Stage 1-4 kernels (src/kernel.cu:436-573 and
 * src/abm.cu:376-1000) and generator assignment (src/abm.cu:567-574).
 * Year steps run it with pg_year_params.model = PG_MODEL_ABM (n_new, year,
 * year_index; the PA and ER fields are ignored).
 */
typedef struct {
    uint32_t size;
    const int32_t* bag;        /* out-degree bag (src/abm.cu:70-88), values in [0, 250] */
    int32_t n_bag;
    const double* rec_prob;    /* p(year difference d) for d in [0, n_rec) (src/abm.cu:91-110) */
    int32_t n_rec;             /* <= 1024; differences outside have p = 0 */
    double alpha;              /* 1-hop share in [0,1]; < 0: uniform per paper */
    double w_pa, w_rec, w_fit; /* >= 0 with a positive sum, or any -1: uniform per paper */
    double fully_random;       /* share of k reserved for uniform papers, [0,1] */
    double same_year;          /* share of a year's papers citing the same year, [0,1] */
} pg_abm_model;

/* Configures the scoring model on ctx (after pg_graph_load or
 * pg_checkpoint_load; kept across years).  Copies the bag and the table. */
int pg_abm_configure(pg_ctx* ctx, const pg_abm_model* model);

/* Per new paper of the last PG_MODEL_ABM year, info[6*i .. 6*i+5]:
 * generator, out-degree, same-year target (-1 if none), 1-hop picks,
 * 2-hop picks, uniform picks. */
int pg_abm_year_info(pg_ctx* ctx, int32_t* info);

/* Candidates keyed by the last PG_MODEL_ABM year's selection:
 * counts[0] = 1-hop entries, counts[1] = 2-hop entries (with repeats). */
int pg_abm_year_counts(pg_ctx* ctx, int64_t* counts);

/* Score components of the last PG_MODEL_ABM year's existing papers:
 * comp[4*c .. 4*c+3] = P, R, F, 0 for c in [0, n_exist). */
int pg_abm_scores(pg_ctx* ctx, float* comp);

/* Static draws of papers [first, first+count) under the configured model
 * (host computation): out-degree, alpha, weights[3*i..3*i+2], fitness peak
 * value; any pointer may be NULL. */
int pg_abm_paper_draws(pg_ctx* ctx, int64_t first, int64_t count, int32_t* out_degree,
                       double* alpha, double* weights, int32_t* fitness);

/* The 1000 cumulative fitness-peak weights f^-1.634 (f = 1..1000) the
 * model draws fitness from (includes/abm.cuh:469-484). */
void pg_abm_fitness_table(double* cdf);

/* Current paper and edge counts. */
int pg_graph_size(const pg_ctx* ctx, int64_t* n_papers, int64_t* n_edges);

/*
 * Copies the graph to caller buffers (any pointer may be NULL):
 * fwd_off/bwd_off [n+1], fwd_adj/bwd_adj [e], in_deg/birth [n],
 * logfit [n].  Rows are sorted ascending.
 */
int pg_graph_download(pg_ctx* ctx, int32_t* fwd_off, int32_t* fwd_adj,
                      int32_t* bwd_off, int32_t* bwd_adj, int32_t* in_deg,
                      int32_t* birth, float* logfit);

/*
 * Copies the last PA year's per-paper draws: cit/kind [n_new*m] in draw
 * order (kind 0 PA, 1 neighbourhood, 2 PA refill; unused slots -1),
 * cnt/near [n_new].  Replaces the per-paper D2H of append_device_to_host
 * (includes/utils.cuh:151-200).  Any pointer may be NULL.
 */
int pg_year_citations(pg_ctx* ctx, int32_t* cit, int8_t* kind, int32_t* cnt,
                      int32_t* near_boundary);

/*
 * Copies the last year's FP32 weights [n_before] and FP64 CDF [n_before]
 * (PA years only).  Any pointer may be NULL.
 */
int pg_year_weights(pg_ctx* ctx, float* w, double* cdf);

/*
 * 2-hop neighbourhoods of `count` roots on the current graph, with the
 * context's cap: out[i*bfs_cap ...] gets the ascending ids, out_len[i] the
 * count.  Exposes the Stage-1 BFS (src/kernel.cu:17-168,
 * src/abm.cu:577-883) on its own for testing and tooling.
 */
int pg_bfs(pg_ctx* ctx, int64_t count, const int32_t* roots, int32_t* out,
           int32_t* out_len);

/*
 * pg_bfs with per-root timing of the device work: ticks[3*i] = wall-clock
 * ticks (hipDeviceAttributeWallClockRate per ms) to the end of the 1-hop
 * level, ticks[3*i+1] = to the end of the query, ticks[3*i+2] = 2-hop
 * candidate entries.  A tooling entry point with no reference counterpart
 * (the reference times whole stages, includes/epoch_profiler.cuh:40-51).
 */
int pg_bfs_timed(pg_ctx* ctx, int64_t count, const int32_t* roots, int32_t* out,
                 int32_t* out_len, int64_t* ticks);

/*
 * pg_bfs without allocation, for callers that cannot allocate device memory
 * (code under HIP stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  roots_dev, out_dev
 * [count*bfs_cap] and out_len_dev [count] are DEVICE buffers; the call only
 * enqueues work (on `stream` when stream_given != 0, NULL being the null
 * stream; else on the context's stream) and returns without waiting.  The
 * neighbourhoods equal pg_bfs's.  A root outside [0, n) gets out_len 0 (the
 * roots are not read on the host).  Scratch is one static device array of
 * about 24 MiB per GPU, held until the process ends: count above 1048576
 * roots is refused (PG_ECAP).  Calls that overlap on one GPU (other streams,
 * other contexts) share that array and race; calls on one stream do not.
 */
int pg_bfs_noalloc(pg_ctx* ctx, int64_t count, const int32_t* roots_dev, int32_t* out_dev,
                   int32_t* out_len_dev, void* stream, int32_t stream_given);

/*
 * Device-side snapshot of the graph state (CSR, degrees, birth years,
 * log-fitness): pg_snapshot_save copies the current state into a second
 * device buffer, pg_snapshot_restore copies it back.  Used to rerun a
 * simulation from the same seed graph without another host upload.
 */
int pg_snapshot_save(pg_ctx* ctx);
int pg_snapshot_restore(pg_ctx* ctx);

/*
 * Year steps completed since pg_graph_load (or in the loaded checkpoint),
 * and the last step's year and year index (-1: none).  Snapshots keep them.
 */
int pg_get_progress(pg_ctx* ctx, int32_t* steps_done, int32_t* last_year,
                    int32_t* last_year_index);

/*
 * Exact year-granular checkpoint and resume.  No reference counterpart (the
 * reference keeps no state between runs).  pg_checkpoint_save writes the
 * graph state -- both CSRs, in-degrees, birth years, the log-fitness of
 * every paper supplied, the Philox seed and the progress above -- plus an
 * optional caller blob (user, user_bytes) to `path` (through a temporary
 * file renamed into place), with an FNV-1a checksum.  pg_checkpoint_load
 * reads it into a context created with the same seed and capacities at
 * least as large; the caller's blob goes to user[0, user_cap) and its size
 * to *user_bytes (nullable).  Log-fitness the context already holds for
 * papers beyond the checkpoint's (pg_graph_load of a longer run of the same
 * seed) is kept, so such a run can resume from a shorter run's checkpoint.  Philox draws depend only on (seed, paper,
 * attempt, year index), so the loaded context continues bit for bit like
 * the one that saved.  The likelihood sequence store is not part of it.
 * PG_EIO on a file error, bad magic, version or checksum; PG_EINVAL on a
 * seed mismatch; PG_ECAP when the state does not fit the context.
 */
int pg_checkpoint_save(pg_ctx* ctx, const char* path, const void* user, int64_t user_bytes);
int pg_checkpoint_load(pg_ctx* ctx, const char* path, void* user, int64_t user_cap,
                       int64_t* user_bytes);

/*
 * Per-kernel-group timing with hipEvents recorded on the context's stream
 * around each group (off by default).  pg_profile_read returns, for each
 * group g < n: the summed device milliseconds, the summed algorithmic HBM
 * bytes and the number of launches; pg_profile_reset clears.  Algorithmic
 * bytes count compulsory traffic only:
 *   WEIGHTS_CDF  20 B per existing paper (12 B read, 8 B FP64 CDF written);
 *   PA_DRAW      8*ceil(log2 N) B per attempt + 4 B per emitted target;
 *   BFS          per distinct root, 16 B + 4 B*deg(v) for the root and each
 *                capped 1-hop paper, + 4 B per output id;
 *   NB_DRAW      20 B per new paper + 8 B per citation slot;
 *   ER           4 B per emitted target + 4 B per new paper;
 *   CSR_APPEND   4*(dE+dN) forward write + 4*(E_old+N_old) read +
 *                4*(E_new+N_new) write + 8 B per new edge (degree RMW);
 *   LIK          4 B per paper-year (in-degree snapshot) + 8 B per paper
 *                (birth, log-fitness) + 4 B per citing paper (count) +
 *                17 B per filled slot (target, kind, the 8 B {log(k+1),
 *                age} record fixed by the sequence, 4 B log-fitness
 *                gathered);
 *   LIK_GRAD     two paper passes: 8 B per paper-year + 44 B per paper
 *                (2x birth and log-fitness, segment offset, per-paper share
 *                sum written and read, dL/dphi written) + 4 B per citing
 *                paper + 33 B per filled slot (17 B as LIK + 8 B share
 *                written + 8 B sorted position and share read back).
 *   LIK_BATCH    per call at B points: shared by the points, 8 B per
 *                paper-year (log(k+1) in both paper passes) + 12 B per paper
 *                (birth twice, segment offset) + 4 B per citing paper + 21 B
 *                per filled slot (target, kind, {log(k+1), age} record,
 *                by-target position, sorted key); per point, 32 B per paper
 *                (log-fitness twice, share sum written and read, dL/dphi) +
 *                20 B per filled slot (log-fitness gathered, share written
 *                and read back).
 */
int pg_profile_enable(pg_ctx* ctx, int on);
int pg_profile_read(pg_ctx* ctx, int n, double* ms, double* bytes, int64_t* launches);
int pg_profile_reset(pg_ctx* ctx);
/* Restricts the timing to the groups whose bit (1u << PG_K_...) is set in
 * mask (default: all).  Each event pair between kernels costs the stream a
 * few microseconds, so a caller timing whole steps may keep only the group
 * it reports.  No reference counterpart (the reference times whole stages). */
int pg_profile_groups(pg_ctx* ctx, uint32_t mask);

/*
 * ---- Citation sequences and the PA likelihood ----
 * No reference counterpart.  A context holds one citation sequence on the
 * device: per year y, the in-degree snapshot of the papers [0, N_y) that
 * exist at the start of the year and the year's PA-model draws (m slots per
 * new paper, in draw order, kinds 0 PA / 1 neighbourhood / 2 PA refill, as
 * pg_year_citations returns them).  Recording fills it from pg_year_step;
 * pg_seq_append_year fills it from host arrays.
 *
 * log L = sum over slots of kind 0 or 2 of  x_t - log(Z_y - P)
 *   x_j = alpha*log(k_j+1) - (year_y - birth_j)/tau + phi_j,  w_j = exp(x_j)
 *   Z_y = sum_{j < N_y} w_j,  P = sum of w over the same paper's earlier slots
 * with x and w in FP32 (the simulator's weight arithmetic, alpha and 1/tau
 * rounded to float) and Z, P, log L and all gradients in FP64.
 */
typedef struct {
    uint32_t size;
    int32_t n_years;
    int32_t m;                 /* slots per new paper */
    int64_t n_papers;          /* papers the sequence refers to (phi length) */
    int64_t paper_years;       /* sum over years of N_y */
    int64_t new_papers;        /* sum over years of the new papers */
} pg_seq_info;

typedef struct {
    uint32_t size;
    double alpha;              /* used when alpha_tau_dev is NULL */
    double tau;                /* years, > 0 */
    const double* alpha_tau_dev; /* optional DEVICE {alpha, tau}: read by the
                                    kernels, so no host round trip */
    const float* logfit_dev;   /* DEVICE float[n_papers]: phi, the parameters */
    void* stream;              /* hipStream_t to run on when stream_given != 0
                                  (NULL is then the null stream) */
    int32_t stream_given;      /* 0: run on the context's own stream */
} pg_lik_params;

typedef struct {
    uint32_t size;
    double* scalars_dev;       /* optional DEVICE double[3]: log L, dL/dalpha,
                                  dL/dtau (the last two 0 for pg_pa_loglik) */
    double* grad_logfit_dev;   /* pg_pa_loglik_grad: DEVICE double[n_papers],
                                  dL/dphi (overwritten) */
    double* host3;             /* optional HOST double[3] as scalars_dev; when
                                  set, the call waits for the result */
} pg_lik_out;

/* Allocates (or re-allocates) and clears the context's sequence store. */
int pg_seq_reserve(pg_ctx* ctx, int32_t m, int64_t max_paper_years,
                   int64_t max_new_papers, int64_t max_papers);
/* on != 0: every later PA pg_year_step appends its year (the in-degree
 * snapshot before the step, the draws after it, the births).  ER years are
 * refused while recording. */
int pg_seq_record(pg_ctx* ctx, int on);
/* Appends one year from host arrays: deg[n_exist], cit/kind[n_new*m],
 * cnt[n_new].  Targets must be in [0, n_exist), kinds in {0,1,2}, counts in
 * [0, m]. */
int pg_seq_append_year(pg_ctx* ctx, int32_t year, int64_t n_exist, int64_t n_new,
                       const int32_t* deg, const int32_t* cit, const int8_t* kind,
                       const int32_t* cnt);
/* Births of papers [0, n_papers) (host array); n_papers must cover every
 * year's N_y + new papers.  Recording sets them itself. */
int pg_seq_set_birth(pg_ctx* ctx, int64_t n_papers, const int32_t* birth);
int pg_seq_get_info(pg_ctx* ctx, pg_seq_info* info);
/* Copies the sequence to host buffers sized from pg_seq_info (any may be
 * NULL): per-year tables, then deg[paper_years], cit/kind[new_papers*m],
 * cnt[new_papers], birth[n_papers]. */
int pg_seq_download(pg_ctx* ctx, int64_t* n_exist, int64_t* n_new, int32_t* year_of,
                    int32_t* deg, int32_t* cit, int8_t* kind, int32_t* cnt,
                    int32_t* birth);
/* log L of the sequence.  Asynchronous on the chosen stream unless
 * out->host3 is set.  Calls on one context must not overlap. */
int pg_pa_loglik(pg_ctx* ctx, const pg_lik_params* p, pg_lik_out* out);
/* log L and its gradient.  The backward paper pass recomputes every weight
 * w_j(y) and the partition sums are recomputed per call (nothing is kept
 * from pg_pa_loglik).  Every sum has a fixed order (dL/dphi via the slots
 * sorted by target, built once per sequence), so repeated calls return the
 * same bits. */
int pg_pa_loglik_grad(pg_ctx* ctx, const pg_lik_params* p, pg_lik_out* out);

/*
 * log L and its gradient at B parameter points in one pass over the
 * sequence (B = 4, 8 or 16): point b has alpha_b, tau_b and its own
 * log-fitness vector.  Each point's results equal pg_pa_loglik_grad's at
 * that point up to FP64 summation order.  The sequence's streams are read
 * once for all points and the per-point arrays interleave the points, so
 * every gathered log-fitness line and every stored share line serves B
 * evaluations.  No reference counterpart (the reference has no likelihood).
 */
typedef struct {
    uint32_t size;
    int32_t n_points;           /* B: 4, 8 or 16 (B * lanes per citing paper <= 64) */
    const double* alpha_tau;    /* HOST double[2B] {alpha_b, tau_b}, used when
                                   alpha_tau_dev is NULL */
    const double* alpha_tau_dev;/* optional DEVICE double[2B] */
    const float* logfit_dev;    /* DEVICE float[n_papers * B]: phi_b of paper j
                                   at j*B + b */
    void* stream;               /* as pg_lik_params */
    int32_t stream_given;
    int32_t arith;              /* PG_LIK_ARITH_SIM or PG_LIK_ARITH_FAST */
} pg_lik_batch_params;

/* Arithmetic of a batched evaluation.  SIM: the weights are the simulator's
 * own (det_expf, the bits the sequence was drawn with), log D and 1/D are
 * the correctly rounded FP64 routines; each point equals pg_pa_loglik_grad
 * up to FP64 summation order.  FAST: the hardware exponential, logarithm
 * and reciprocal (weights within a few ulp, log D within about 1e-7, 1/D
 * within an ulp after one Newton step); the partition sums stay FP64.  FAST
 * agrees with pg_pa_loglik_grad to about 1e-7 relative and is bound by
 * memory traffic rather than by the weight arithmetic. */
#define PG_LIK_ARITH_SIM 0
#define PG_LIK_ARITH_FAST 1

typedef struct {
    uint32_t size;
    double* scalars_dev;        /* optional DEVICE double[3B]: {log L, dL/dalpha,
                                   dL/dtau} of point b at 3b */
    double* grad_logfit_dev;    /* DEVICE double[n_papers * B]: dL/dphi_b of
                                   paper j at j*B + b (overwritten) */
    double* host;               /* optional HOST double[3B] as scalars_dev; when
                                   set, the call waits for the result */
} pg_lik_batch_out;

int pg_pa_loglik_grad_batch(pg_ctx* ctx, const pg_lik_batch_params* p, pg_lik_batch_out* out);

/*
 * ---- Batched ensembles ----
 * No reference counterpart (the reference runs one simulation per process).
 * Many independent PA simulations -- each with its own seed graph,
 * parameters and Philox key -- advance together: every kernel of a year step
 * covers all of them, and a whole run (all years) is captured once as a
 * hipGraph and replayed.  Simulation s reproduces, bit for bit, a plain
 * context loaded with the same seed graph and stepped with the same
 * parameters (year = first_year + y, year_index = y).
 */
typedef struct pg_ensemble pg_ensemble;

typedef struct {
    uint32_t size;
    int32_t device;
    int32_t n_sims;
    int32_t n_years;           /* growth years per run */
    int32_t m;                 /* citations per new paper (every simulation) */
    int32_t bfs_cap;           /* 2-hop cap (<= 4096) */
    int32_t first_year;        /* calendar year of growth year 0 */
    double growth_rate;        /* new papers per year: ceil(N * rate) */
} pg_ensemble_config;

typedef struct {
    uint32_t size;
    int32_t n_nb;              /* neighbourhood draws per paper (<= m-1) */
    int64_t n0;                /* seed papers */
    int64_t seed_edges;        /* bound on the distinct seed citations loaded */
    float alpha;               /* in-degree exponent */
    float inv_tau;             /* 1/tau in 1/years */
    uint64_t seed;             /* Philox key */
} pg_sim_params;

typedef struct {
    uint32_t size;
    int64_t edges_added;       /* over all simulations and years */
    int64_t near_boundary;
    int64_t shortfall;
    int64_t pa_attempts;
    double device_ms;          /* GPU time of the run (hipEvents) */
    double bfs_ms;             /* pg_ensemble_step: GPU time of the 2-hop kernels
                                  (hipEvents around them); 0 for pg_ensemble_run */
    double bfs_bytes;          /* algorithmic bytes of the 2-hop kernels (the
                                  BFS formula of pg_profile_read) */
} pg_ensemble_stats;

/* Launch shapes that change speed, never results. */
typedef struct {
    uint32_t size;
    int32_t pa_lanes;          /* PA attempts evaluated together per paper:
                                  8, 16, 32, 64 (0 = by citations per paper) */
    int32_t lik_lanes;         /* likelihood slot pass: lanes per citing paper
                                  1, 2, 4, 8 with 8 * lanes >= citations per
                                  paper (0 = the fewest such) */
    int32_t bfs_mode;          /* workgroup 2-hop path for candidate lists
                                  above 4096 entries: 0 = radix-sorted in LDS
                                  when the entries below the answer bound
                                  (or a histogram cut) fit 7168, else staged
                                  in id buckets up to 65536 entries, else
                                  swept; 1 = always swept; 2 = staged
                                  whenever the list fits a staging slot
                                  (lists up to 4096 entries: radix-sorted) */
    int32_t bfs_small_max;     /* waves per 2-hop query: the most candidate
                                  entries a root may bring to the one-wavefront
                                  small tier, 256/512/1024/2048 (0 = 2048) */
    int32_t bfs_mid_max;       /* ... to the one-wavefront medium tier,
                                  512/1024/2048/4096 (0 = 4096); roots past
                                  it take a 512-thread workgroup */
    int32_t merge_out;         /* backward-CSR merge tile: output positions
                                  per block, 256/512/1024/2048 (0 = 1024) */
} pg_tuning;

int pg_ensemble_create(const pg_ensemble_config* cfg, const pg_sim_params* sims,
                       pg_ensemble** out);
void pg_ensemble_destroy(pg_ensemble* ens);
/* Error text of the last failed call on ens (NULL: of pg_ensemble_create). */
const char* pg_ensemble_error(const pg_ensemble* ens);
/* Seed graph of simulation `sim` (loaded in order 0, 1, ...): n == n0
 * papers, logfit for all of its final papers; after the last one the seed
 * state is kept for pg_ensemble_restore. */
int pg_ensemble_load(pg_ensemble* ens, int32_t sim, int64_t n, const int32_t* birth,
                     const float* logfit, int64_t n_edges, const int32_t* src,
                     const int32_t* dst);
/* Instead of pg_ensemble_load: every simulation's seed graph and log-fitness
 * generated on the device, bit for bit what pg_seed_graph(key, n0, year_lo,
 * year_hi, mean_cites) and pg_seed_logfit(key, 0, final papers, mu, sigma)
 * give for the simulation's Philox key (its pg_sim_params.seed).  seed_edges
 * must bound each simulation's draws (PG_ECAP otherwise).  No reference
 * counterpart (the reference reads its seed graph from files). */
int pg_ensemble_generate(pg_ensemble* ens, int32_t year_lo, int32_t year_hi, double mean_cites,
                         double mu, double sigma);
/* All years for all simulations, from the seed state.  use_graph != 0:
 * captured once into a hipGraph, replayed on later runs. */
int pg_ensemble_run(pg_ensemble* ens, int32_t use_graph, pg_ensemble_stats* stats);
int pg_ensemble_restore(pg_ensemble* ens);
/* One growth year (the next one) for all simulations, eagerly, from the seed
 * state or after earlier steps; the same kernels as pg_ensemble_run, which
 * needs the seed state.  Lets a caller inspect every year. */
int pg_ensemble_step(pg_ensemble* ens, pg_ensemble_stats* stats);
/* The last year's draws of simulation `sim` in its local ids, as
 * pg_year_citations returns them (cit/kind [n_new*m], cnt/near [n_new]), and
 * its FP32 weights and FP64 CDF [n_exist] of that year.  Any pointer may be
 * NULL. */
int pg_ensemble_year_draws(pg_ensemble* ens, int32_t sim, int32_t* cit, int8_t* kind,
                           int32_t* cnt, int32_t* near_boundary, float* w, double* cdf);
int pg_ensemble_sim_size(pg_ensemble* ens, int32_t sim, int64_t* n_papers, int64_t* n_edges);
/* One simulation's CSR in its local ids (buffers sized from sim_size: the
 * papers after the years done so far). */
int pg_ensemble_download(pg_ensemble* ens, int32_t sim, int32_t* fwd_off, int32_t* fwd_adj,
                         int32_t* bwd_off, int32_t* bwd_adj, int32_t* in_deg,
                         int32_t* birth);
int pg_ensemble_set_tuning(pg_ensemble* ens, const pg_tuning* t);
/* The same launch shapes for a plain context (pg_year_pa, pg_bfs,
 * pg_pa_loglik[_grad]).  No reference counterpart: the reference fixes its
 * launch shapes at compile time (src/kernel.cu:1600-1640). */
int pg_set_tuning(pg_ctx* ctx, const pg_tuning* t);

/*
 * ---- Autotuner ----
 * No reference counterpart (the reference fixes its launch shapes at compile
 * time, src/kernel.cu:1600-1640).  Times candidate launch shapes on the
 * caller's own workload with hipEvents -- one untimed run per candidate
 * (graph capture, table rebuilds), then `reps` timed runs, median -- and
 * keeps the fastest.
 */
#define PG_TUNE_MAX_TRIALS 16
typedef struct {
    uint32_t size;
    int32_t n_trials;
    pg_tuning trial[PG_TUNE_MAX_TRIALS];
    double trial_ms[PG_TUNE_MAX_TRIALS];  /* median device ms per candidate */
    pg_tuning chosen;                      /* now set on the ensemble/context */
    double chosen_ms;
    double default_ms;                     /* the default shapes' median */
    double max_rel_dev;                    /* pg_lik_autotune: largest relative
                                              difference of log L, dL/dalpha,
                                              dL/dtau from the default shapes'
                                              (summation order); 0 for ensembles */
} pg_tune_report;

/* Ensemble: PA lanes per paper (8, 16, 32, 64) with the default 2-hop mode,
 * then the 2-hop modes (0, 1, 2) with the best lanes; every trial is a whole
 * graph-captured run from the seed state.  Candidates must reproduce the
 * same edge, attempt and near-boundary counts (PG_ESTATE otherwise).  Leaves
 * the ensemble in its seed state with the chosen shapes. */
int pg_ensemble_autotune(pg_ensemble* ens, int32_t reps, pg_tune_report* report);
/* Likelihood: lanes per citing paper in the slot pass (1, 2, 4, 8 with
 * 8 * lanes >= m), each trial one pg_pa_loglik_grad at p's parameters on the
 * context's stream.  Leaves the context with the chosen shape. */
int pg_lik_autotune(pg_ctx* ctx, const pg_lik_params* p, int32_t reps, pg_tune_report* report);
/* Year steps of a plain context, on the caller's schedule years[0..n_years)
 * run from the context's current state: PA lanes (8, 16, 32, 64), then the
 * 2-hop modes, the wavefront tiers' candidate limits (small 512, 1024; medium
 * 2048) and the merge tile (512, 2048), each step keeping the best so far.
 * Every trial restores the state (pg_snapshot_save is called first and
 * replaces the context's snapshot) and must reproduce the same edge,
 * attempt and near-boundary counts (PG_ESTATE otherwise).  Leaves the
 * context in the state it was called in, with the chosen shapes. */
int pg_year_autotune(pg_ctx* ctx, const pg_year_params* years, int32_t n_years, int32_t reps,
                     pg_tune_report* report);

/* Ceil growth rule of the reference (src/abm.cu:233-240,
 * src/kernel.cu:1690): ceil(n * rate) in double precision. */
int64_t pg_growth_count(int64_t n, double rate);

/*
 * Synthetic seed graph (host, OpenMP, Philox-keyed so every run matches):
 * n papers with birth years uniform in [year_lo, year_hi], ids sorted by
 * birth; each cites min(Poisson(mean_cites), #older) distinct older papers
 * chosen uniformly.  birth[n]; src/dst sized by *n_edges on return: call once
 * with src == NULL to get the count.  Log-fitness is drawn separately by
 * pg_seed_logfit.
 */
int pg_seed_graph(uint64_t seed, int64_t n, int32_t year_lo, int32_t year_hi,
                  double mean_cites, int32_t* birth, int64_t* n_edges,
                  int32_t* src, int32_t* dst);

/* log-fitness phi_i ~ Normal(mu, sigma^2) for papers [first, first+count). */
int pg_seed_logfit(uint64_t seed, int64_t first, int64_t count, double mu,
                   double sigma, float* out);

#ifdef __cplusplus
}
#endif
#endif
