// pg_ctx.h -- the context object behind the C ABI (private to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_internal.h"

// A citation sequence held on the device for the PA likelihood: per year,
// the in-degree snapshot at the start of the year and the year's draws
// (m slots per new paper, in draw order).
struct pg_seq_store {
    bool reserved = false;
    int32_t m = 0;
    int64_t cap_paper_years = 0, cap_new = 0, cap_papers = 0;
    // host copies of the per-year tables
    std::vector<int64_t> n_exist, n_new, deg_off, new_off;
    std::vector<int32_t> year_of;
    int64_t n_papers = 0;        // papers in the sequence (ids 0..n_papers-1)
    int64_t paper_years = 0;     // sum of n_exist
    int64_t new_total = 0;       // sum of n_new
    int64_t filled = 0;          // filled slots (sum of cnt)
    // the year pg_year_step has snapshotted and whose draws follow
    bool pending = false;
    int32_t pending_year = 0;
    int64_t pending_n_exist = 0, pending_n_new = 0;
    // device arrays
    int32_t* deg = nullptr;      // [cap_paper_years] in-degree snapshots
    int32_t* cit = nullptr;      // [cap_new * m] targets, draw order
    int8_t* kind = nullptr;      // [cap_new * m] 0 PA, 1 neighbourhood, 2 PA refill
    int32_t* cnt = nullptr;      // [cap_new] filled slots per new paper
    int32_t* birth = nullptr;    // [cap_papers]
    // device copies of the per-year tables and the launch tables
    bool tables_dirty = true;
    int64_t* d_n_exist = nullptr;
    int64_t* d_n_new = nullptr;
    int64_t* d_deg_off = nullptr;
    int64_t* d_new_off = nullptr;
    int32_t* d_year_of = nullptr;
    int32_t* d_ev_block = nullptr;   // event block -> year, then first new paper
    int32_t* d_ev_year_blk = nullptr;// year -> first event block (n_years+1)
    int64_t n_ev_blocks = 0;
    pg::LikCiteShape cite_shape{1, 4};  // slot-pass lanes per paper and slots per lane
    int64_t n_paper_blocks = 0;
    double* part = nullptr;          // per-block partial sums
    int64_t part_cap = 0;
    double* year_sums = nullptr;     // per year: Z, A
    int64_t table_cap_years = 0;
    int64_t table_cap_blocks = 0;
    double* d_scalars = nullptr;     // loglik, d/dalpha, d/dtau
    // slots sorted by target (built when the tables are rebuilt)
    uint32_t* skey = nullptr;        // [new * m] slot targets, sorted
    uint32_t* pos_t = nullptr;       // [new * m] by-target position, slot-major
    int32_t* cit_t = nullptr;        // [new * m] targets, slot-major
    int8_t* kind_t = nullptr;        // [new * m] kinds, slot-major
    float2* la_t = nullptr;          // [new * m] {log(k+1), age} of each slot's target, slot-major
    float* Lpy = nullptr;            // [paper_years] log(k+1) of every in-degree snapshot
    int32_t* seg_off = nullptr;      // [cap_papers + 1]
    double* share = nullptr;         // [filled] per-slot dL/dphi shares, by-target order
    double* evt = nullptr;           // [cap_papers]
    double* carry = nullptr;         // [2 * chunks]
    int64_t carry_cap = 0;
    // batched evaluation at bB points (pg_pa_loglik_grad_batch); built for
    // the sequence of size b_key = {years, citing papers, filled slots, papers}
    int32_t bB = 0;
    int64_t b_key[4] = {-1, -1, -1, -1};
    pg::LikCiteShape b_shape{1, 4};
    int64_t b_nE = 0, b_nP = 0;
    int32_t* b_ev_block = nullptr;   // slot-pass block table of the batched shape
    int32_t* b_year_blk = nullptr;
    double* b_part = nullptr;
    double* b_year_sums = nullptr;
    double* b_share = nullptr;       // [filled * bB]
    double* b_evt = nullptr;         // [n_papers * bB]
    int64_t* b_job_start = nullptr;  // segment-sum jobs (cut at paper boundaries)
    int32_t* b_job_tag = nullptr;
    int64_t b_n_jobs = 0;
    double* b_bigpart = nullptr;     // [pieces * bB]
    int32_t* b_big_paper = nullptr;
    int32_t* b_big_first = nullptr;
    int32_t b_n_big = 0;
    double* b_at = nullptr;          // [2 * bB] alpha/tau copied from the host
    double* b_out = nullptr;         // [3 * bB]
};

// The reference's scoring model (pg_abm.hip): configuration, its device
// copies, and the per-year buffers.
struct pg_abm_store {
    bool on = false;
    int32_t n_bag = 0, n_rec = 0;
    double alpha = 0, w_pa = 0, w_rec = 0, w_fit = 0, frc = 0, syp = 0;
    std::vector<int32_t> h_bag;     // host copy (pg_abm_paper_draws)
    std::vector<double> h_cdf;      // fitness table
    int32_t* bag = nullptr;
    double* rec = nullptr;          // [kAbmRecMax]
    double* cdf = nullptr;          // [kAbmFitMax]
    int16_t* fit = nullptr;         // [max_papers] fitness peak values
    int64_t fit_n = 0;              // fit holds ids [0, fit_n)
    float4* comp = nullptr;         // [max_papers] P, R, F of the year
    uint64_t* parts = nullptr;      // [3 * kAbmMaxParts]
    int32_t* cnt = nullptr;         // [kAbmRecMax + 1]: recency counts, then the 2^21 flag
    float* rt = nullptr;            // [kAbmRecMax]
    double* tot = nullptr;          // [2]
    int32_t* syt = nullptr;         // [max_new] same-year targets
    uint32_t* sy_keys = nullptr;    // [max_new]
    uint32_t* sy_vals = nullptr;    // [max_new]
    int32_t* sel = nullptr;         // [2 * max_new * kAbmSlots] selected ids per paper and level
    int32_t* nsel = nullptr;        // [2 * max_new]
    int32_t* info = nullptr;        // [6 * max_new]
    unsigned long long* cand = nullptr;  // [2] candidates keyed per level this year
    int32_t* kc = nullptr;          // [2 * max_new] candidates keyed per paper and level
    int64_t last_n = -1, last_d = 0;  // the last year's existing and new papers
    int64_t last_cand[2] = {0, 0};
};

struct pg_ctx {
    pg_config cfg{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bool loaded = false;
    int64_t n = 0, e = 0, n_logfit = 0;
    bool logfit_owned = false;        // logfit[0, n_logfit) holds supplied values
    int32_t* birth = nullptr;
    int32_t* in_deg = nullptr;
    float* logfit = nullptr;
    int32_t* fwd_off = nullptr;
    int32_t* fwd_adj = nullptr;
    int32_t* bwd_off[2] = {nullptr, nullptr};
    int32_t* bwd_adj[2] = {nullptr, nullptr};
    int cur = 0;
    float* w = nullptr;
    double* cdf = nullptr;
    int32_t* cdf_idx = nullptr;  // kCdfIdx + 2 buckets of the current CDF
    double* scan_agg = nullptr;
    int32_t* cit = nullptr;
    int8_t* kind = nullptr;
    int32_t* cnt = nullptr;
    int32_t* near = nullptr;
    int32_t* anext = nullptr;
    int64_t* new_off = nullptr;
    int64_t year_edge_cap = 0;
    int32_t* src_of = nullptr;
    uint32_t* k_sorted = nullptr;
    uint32_t* v_sorted = nullptr;
    uint32_t* k_tmp = nullptr;
    uint32_t* v_tmp = nullptr;
    int64_t* scratch = nullptr;
    unsigned long long* d_sum = nullptr;
    unsigned long long* h_sum = nullptr;  // pinned host copy of d_sum
    int32_t* d_tmp_i32 = nullptr;
    int64_t d_tmp_cap = 0;
    int64_t last_n_new = 0, last_n_exist = 0;
    int last_m = 0, last_model = -1;
    // per-kernel-group profiling
    bool prof = false;
    uint32_t prof_mask = 0xFFFFFFFFu;  // groups timed while prof is on
    hipEvent_t pev[PG_K_COUNT][2] = {};
    hipEvent_t pev2[PG_K_COUNT][2] = {};
    double prof_ms[PG_K_COUNT] = {};
    double prof_bytes[PG_K_COUNT] = {};
    int64_t prof_calls[PG_K_COUNT] = {};
    // device snapshot of the graph state
    bool snap = false;
    int64_t snap_n = 0, snap_e = 0;
    // year steps completed since the graph was loaded, and the last one's
    // year and year index (kept by snapshots and checkpoints)
    int32_t steps_done = 0, last_year = 0, last_year_index = -1;
    int32_t snap_steps = 0, snap_last_year = 0, snap_last_year_index = -1;
    int32_t* s_birth = nullptr;
    int32_t* s_in_deg = nullptr;
    int32_t* s_fwd_off = nullptr;
    int32_t* s_fwd_adj = nullptr;
    int32_t* s_bwd_off = nullptr;
    int32_t* s_bwd_adj = nullptr;
    // neighbourhoods of the year's distinct roots
    int32_t* S_buf = nullptr;
    uint64_t* Q_buf = nullptr;
    int32_t* Q_idx = nullptr;  // kQIdx bucket starts of each root's Q
    int32_t* ns_buf = nullptr;
    unsigned long long* byte_slots = nullptr;  // [kByteSlots * kByteSlotStride] 2-hop byte counts, zero between launches
    uint32_t* root_bits = nullptr; // [max_papers/32] roots of the year, zero between years
    int32_t* r_wcount = nullptr;   // [max_papers/32] popcounts of root_bits
    int64_t* r_wprefix = nullptr;  // [max_papers/32 + 1] their exclusive prefix
    int64_t* r_excl = nullptr;     // [max_new + 1]: [max_new] = distinct roots
    int32_t* root_slot = nullptr;
    int32_t* uroot = nullptr;
    int32_t* big_list = nullptr;
    pg::BfsStage stage;           // staging pool of the workgroup 2-hop path
    // launch shapes (pg_set_tuning): 0 = default
    int tune_pa_lanes = 0;
    int tune_lik_lanes = 0;
    int tune_bfs_mode = 0;
    int tune_small_max = 0;
    int tune_mid_max = 0;
    int tune_merge_out = 0;
    // citation-sequence store for the likelihood; recording appends each PA year
    pg_seq_store seq;
    bool seq_recording = false;
    pg_abm_store abm;
};

// Host helpers shared by the ABI source files.
#define PG_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);               \
            return PG_EHIP;                                                             \
        }                                                                               \
    } while (0)

inline int fail(pg_ctx* ctx, int code, const std::string& msg) {
    ctx->err = msg;
    return code;
}

// Device allocation of `count` elements (at least one).  Test hook: with
// PG_TEST_FAIL_ALLOC_BYTES=N in the environment, every request of N bytes
// or more fails as out of memory, so tests can check that a failed growth
// leaves a context consistent.
template <typename T>
inline hipError_t dalloc(T** p, int64_t count) {
    if (count <= 0) count = 1;
    const size_t bytes = (size_t)count * sizeof(T);
    if (const char* lim = std::getenv("PG_TEST_FAIL_ALLOC_BYTES")) {
        const long long n = std::atoll(lim);
        if (n > 0 && bytes >= (size_t)n) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}

// Per-year edge buffers (src_of, radix buffers, scratch) for `need` edges.
int ensure_year_edges(pg_ctx* ctx, int64_t need);
// Radix key bits covering ids below n.
int key_bits_for(int64_t n);
// The current CSR as kernels read it.
pg::CsrView view(const pg_ctx* c);

// Why a pg_tuning is invalid (nullptr: valid), shared by the set_tuning calls.
const char* tuning_error(const pg_tuning* t);

// Neighbourhood buffers sized by cfg.max_new_per_year and cfg.bfs_cap.
int ensure_nb_buffers(pg_ctx* ctx);
// Staging pool of the workgroup 2-hop path (allocated once per context).
int ensure_bfs_stage(pg_ctx* ctx);

// Looks up the device address of pg_bfs_noalloc's static scratch on the
// context's device (current), once per device (pg_noalloc.hip).
int noalloc_setup(pg_ctx* ctx);

// Scoring model (pg_abm.hip): frees its device memory; runs one year.
void abm_release(pg_ctx* ctx);
int abm_year_step(pg_ctx* ctx, const pg_year_params* p, pg_year_stats* stats);

// Sequence store (pg_lik.hip): frees its device memory; appends the PA year
// that pg_year_step is about to run (in-degree snapshot) and has just drawn.
void seq_release(pg_ctx* ctx);
int seq_record_snapshot(pg_ctx* ctx, int32_t year, int64_t n_exist, int64_t n_new, int32_t m,
                        int model);
int seq_record_draws(pg_ctx* ctx, int64_t filled);
