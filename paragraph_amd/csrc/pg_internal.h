// pg_internal.h -- kernel launchers and device helpers shared by the HIP
// translation units of libparagraph.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pg_math.h"

namespace pg {

constexpr int kTile = 4096;          // elements per scan / radix tile
constexpr int kBfsCapMax = 4096;     // largest neighbourhood cap
constexpr int kMaxCites = 64;        // largest citations per paper
constexpr int kAttemptsPerCite = 64; // PA / neighbourhood attempt budget
constexpr double kNearRel = 1e-12;   // boundary-case tolerance (relative)

// First index k in [0,n) with cdf[k] > t, clamped to n-1.
__device__ __forceinline__ int64_t cdf_upper_bound(const double* __restrict__ cdf, int64_t n, double t) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

// Orders this wavefront's LDS accesses before the barrier against its
// accesses after it, across lanes (a compiler barrier; the LDS serves one
// wavefront's instructions in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bucket index of a root's fixed-point neighbourhood CDF Q[0,ns) (total Z):
// bucket of a value x is x >> q_shift(Z) (at most 63), monotone in x, and
// qidx[k] (k <= kQBuckets) is the first s whose bucket is >= k, ns if none;
// the first s with Q[s] > t then lies in [qidx[key(t)], qidx[key(t) + 1]].
constexpr int kQBuckets = 64;
// Spread counters for the 2-hop kernels' byte counts (u64, kByteSlotStride
// apart: one cache line each).
constexpr int kByteSlots = 64;
constexpr int kByteSlotStride = 16;
constexpr int kQIdx = kQBuckets + 1;  // int32 per root

__device__ __forceinline__ int q_shift(uint64_t Z) {
    const int bl = Z ? 64 - __clzll((long long)Z) : 0;
    return bl > 6 ? bl - 6 : 0;
}

// Index entries of element s (inclusive prefix qs, previous prefix qp) of
// a root's CDF of ns elements.
__device__ __forceinline__ void q_index_put(int32_t* __restrict__ qidx, int s, int ns, uint64_t qp,
                                            uint64_t qs, int sh) {
    const int k0 = s == 0 ? 0 : (int)(qp >> sh) + 1;
    const int k1 = (int)(qs >> sh);
    for (int k = k0; k <= k1; ++k) qidx[k] = s;
    if (s == ns - 1)
        for (int k = k1 + 1; k <= kQBuckets; ++k) qidx[k] = ns;
}

// CDF bucket index: bucket of a CDF value or target t is
// floor(t * kCdfIdx / Z) clamped to [0, kCdfIdx], monotone in t (one
// rounding of a product by a positive constant), and idx[k] is the first
// paper whose bucket is >= k (idx[kCdfIdx + 1] = n).  Every paper below
// idx[key(t)] has cdf <= t and paper idx[key(t) + 1] has cdf > t, so the
// first paper with cdf > t lies in [idx[key(t)], idx[key(t) + 1]]: a
// search there returns what the full binary search returns.
constexpr int kCdfIdx = 1 << 18;  // 4 to 10 papers per bucket at 2-2.7M papers

__device__ __forceinline__ double cdf_key_scale(double Z) {
    return Z > 0.0 ? (double)kCdfIdx / Z : 0.0;
}

__device__ __forceinline__ int cdf_key(double t, double kscale) {
    const double x = t * kscale;
    return x <= 0.0 ? 0 : (x >= (double)kCdfIdx ? kCdfIdx : (int)x);
}

// cdf_upper_bound through the index when idx is non-null.
__device__ __forceinline__ int64_t cdf_search(const double* __restrict__ cdf,
                                              const int32_t* __restrict__ idx, int64_t n, double t,
                                              double kscale) {
    if (!idx) return cdf_upper_bound(cdf, n, t);
    const int k = cdf_key(t, kscale);
    int64_t lo = idx[k], hi = idx[k + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

__device__ __forceinline__ bool cdf_near(const double* __restrict__ cdf, int64_t n, int64_t k,
                                         double t, double eps) {
    return (k > 0 && t - cdf[k - 1] <= eps) || (k < n - 1 && cdf[k] - t <= eps);
}

// ---- batched simulations (ensemble) ----
// Every simulation of an ensemble owns a region of one global paper-id
// space starting at `base` (a multiple of kTile, so its CDF tiles and bits
// are those of a plain run); ids are local + base.  One BatchSim per
// simulation and year, in region order.
struct BatchSim {
    int64_t base;          // global id of the simulation's paper 0
    int64_t n_exist;       // papers at the start of the year
    int64_t new_first;     // its first new paper in the year's batch order
    int64_t n_new;
    int64_t tile_first;    // its first CDF tile in the year's tile order
    uint64_t seed;         // Philox key
    float alpha, inv_tau;
    int32_t n_pa, n_nb;
};

struct Batch {
    const BatchSim* sims;  // device table [n_sims]
    int32_t n_sims;
    int32_t year;
    int64_t n_tiles;       // CDF tiles of the year over all simulations
    int64_t n_new;         // new papers of the year over all simulations
};

// Simulation of batch-order new paper p / tile t (binary search).
__device__ __forceinline__ int batch_sim_of_new(const Batch& b, int64_t p) {
    int lo = 0, hi = b.n_sims - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.sims[mid].new_first <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int batch_sim_of_tile(const Batch& b, int64_t t) {
    int lo = 0, hi = b.n_sims - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.sims[mid].tile_first <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---- launchers (stream-ordered, no host synchronisation) ----

// w[i] = pa_weight(...) for i < n; cdf[i] = FP64 inclusive prefix of w.
// tile_scratch: ceil(n/4096) doubles.
// (cdf_idx: kCdfIdx + 2 int32, the bucket index of the new CDF; nullable)
void launch_weights_cdf(hipStream_t s, int64_t n, const int32_t* in_deg, const int32_t* birth,
                        const float* logfit, int32_t year, float alpha, float inv_tau,
                        float* w, double* cdf, double* tile_scratch, int32_t* cdf_idx);

// Batched form: per simulation of the year, w and the FP64 CDF over its
// papers [base, base + n_exist), with its own alpha/inv_tau; tile_scratch:
// b.n_tiles doubles.
void launch_weights_cdf_batch(hipStream_t s, const Batch& b, const int32_t* in_deg,
                              const int32_t* birth, const float* logfit, float* w, double* cdf,
                              double* tile_scratch);

// Exclusive prefix sum of int32 in[0,n) into int64 out[0,n]; out[n] = total.
// scratch: at least scan_scratch_elems(n) int64.
int64_t scan_scratch_elems(int64_t n);
void launch_excl_scan_i32(hipStream_t s, const int32_t* in, int64_t n, int64_t* out, int64_t* scratch);

// PA phase 1: n_pa distinct draws per new paper from the global CDF.
// lanes: attempts evaluated together per paper (8/16/32/64; 0 = by n_pa).
int pa_draw_lanes(int n_pa);
void launch_pa_draw(hipStream_t s, int64_t n_exist, const double* cdf, const int32_t* cdf_idx,
                    int64_t n_new, int m, int n_pa, uint64_t seed, uint32_t year_index,
                    int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* near, int32_t* anext,
                    int lanes = 0);
// Batched: every simulation's new papers of the year (cdf: the global FP64
// CDF array; targets are written as global ids).
void launch_pa_draw_batch(hipStream_t s, const Batch& b, int m, uint32_t year_index, int lanes,
                          const double* cdf, int32_t* cit, int8_t* kind, int32_t* cnt,
                          int32_t* near, int32_t* anext);

struct CsrView {
    const int32_t* fwd_off;
    const int32_t* fwd_adj;
    const int32_t* bwd_off;
    const int32_t* bwd_adj;
    int64_t n;
};

// ---- the reference's scoring model (k_abm.hip) ----
constexpr int kAbmRecMax = 1024;   // recency table entries (year differences 0..1023)
constexpr int kAbmMaxParts = 2048; // blocks of the score sums
constexpr int kAbmSlots = 256;     // per paper and level: selected ids (at most 249)

// One year of the model: N existing papers, D new; the bag and the
// per-paper parameter rules of the configured model.
struct AbmYear {
    uint64_t seed;
    int64_t N, D;
    uint32_t yi;
    int32_t year;
    const int32_t* bag;
    int32_t n_bag;
    double alpha, w_pa, w_rec, w_fit, frc;
};

void launch_abm_fit(hipStream_t s, int64_t lo, int64_t hi, uint64_t seed, const double* cdf,
                    int16_t* fit);
int abm_sum_parts(int64_t N);
void launch_abm_scores(hipStream_t s, int64_t N, const int32_t* in_deg, const int32_t* birth,
                       const int16_t* fit, int32_t year, int32_t n_rec, const double* prob,
                       uint64_t* parts, int32_t* cnt, int32_t* flag, double* tot, float* rt,
                       float4* comp);
void launch_abm_same_keys(hipStream_t s, int64_t N, int64_t D, uint64_t seed, uint32_t yi,
                          uint32_t* keys, uint32_t* vals, int32_t* syt);
void launch_abm_same_mark(hipStream_t s, int64_t N, int64_t D, int64_t K, uint64_t seed,
                          uint32_t yi, const uint32_t* order, int32_t* syt);
// kc[l], kc[D + l]: candidate entries keyed per level for paper l (with repeats).
void launch_abm_select(hipStream_t s, const AbmYear& y, const CsrView& g, const float4* comp,
                       const int32_t* syt, int32_t* sel, int32_t* nsel, int32_t* kc);
void launch_abm_count(hipStream_t s, const AbmYear& y, const int32_t* syt, const int32_t* nsel,
                      int32_t* cnt, int32_t* info);
void launch_abm_write(hipStream_t s, const AbmYear& y, const int32_t* syt, const int32_t* sel,
                      const int32_t* info, const int64_t* new_off, int64_t fwd_base,
                      int32_t* fwd_off, int32_t* fwd_adj, int32_t* src_of);

// Staging pool of the workgroup 2-hop path: n_slots slots of cap_entries
// int32 ids in buf; busy[k] is 0 while slot k is free (zeroed once at
// allocation; each workgroup that takes a slot clears its flag when done).
// A workgroup holds at most one slot, so n_slots above the number of
// resident workgroups always leaves one free; a workgroup that finds none,
// or whose candidate list does not fit, takes the unstaged sweep.
struct BfsStage {
    int32_t* buf = nullptr;
    int32_t* busy = nullptr;
    int32_t n_slots = 0;
    int32_t cap_entries = 0;
    int64_t max_total = 0;   // larger candidate lists (row entries) are swept
    int32_t radix = 0;       // 1: every list tries the in-LDS radix sort first
    // one-wavefront tiers: the most 2-hop candidates a root may bring to the
    // small and to the medium tier (0: the tier's LDS capacity); roots past
    // them go on to the next tier.  Speed only: every tier builds the same S.
    int32_t small_max = 0;
    int32_t mid_max = 0;
};
// Slots for one device: twice the resident workgroups of the path.
int bfs_stage_slots(int device);
constexpr int32_t kBfsStageEntries = 1 << 17;
// Default bound on the candidate entries a root stages: above it the sweep
// was faster on the 2M-paper citation graphs (few long hub rows, where the
// sweep reads only the rows' heads).
constexpr int64_t kBfsStageDefaultMax = 1 << 16;
// The pool view that a launch gets for tuning mode m (0 default bound,
// 1 never staged, 2 staged whenever the list fits a slot).
BfsStage bfs_stage_view(const BfsStage& pool, int mode, int small_max = 0, int mid_max = 0);
// Candidate limits a tuning may give the wavefront tiers (their LDS windows).
constexpr int kBfsSmallWinMax = 2048;
constexpr int kBfsMidWinMax = 4096;

// PA phases 2-3.  Roots (first PA pick of each paper) are deduplicated and
// numbered in ascending id order: root_slot[i] = distinct-root index u of
// paper i (-1 if none), uroot[u] = root id, *n_unique = number of distinct
// roots (on the device).  root_bits (ceil(n_exist/32) words) must be zero on
// entry and is again on exit; wcount/wprefix: ceil(n_exist/32) int32 and
// that + 1 int64; scratch: scan scratch for that many.
void launch_root_dedup(hipStream_t s, int64_t n_exist, int64_t n_new, int m, const int32_t* cit,
                       const int32_t* cnt, uint32_t* root_bits, int32_t* wcount, int64_t* wprefix,
                       int64_t* scratch, int64_t* n_unique, int32_t* root_slot, int32_t* uroot);
// 2-hop neighbourhood S (S_buf[u*cap...], ns_buf[u]) and its fixed-point
// CDF (Q_buf[u*cap...]) of every distinct root: a wavefront per small or
// medium root, a workgroup per large one (big_list: scratch of 2*max_unique;
// big_count[0..1] = 0 on entry); grids of max_unique blocks.
void launch_bfs_root(hipStream_t s, const CsrView& g, const float* w, const int32_t* uroot,
                     const int64_t* n_unique, int64_t max_unique, int cap, int32_t* S_buf,
                     uint64_t* Q_buf, int32_t* q_idx, int32_t* ns_buf, int32_t* big_list,
                     int32_t* big_count,
                     unsigned long long* bytes_out, const BfsStage& stage,
                     unsigned long long* byte_slots);
// Per paper: neighbourhood draws from its root's S, then PA refill.
void launch_nb_draw(hipStream_t s, int64_t n_exist, int64_t n_new, int m, int n_nb, int cap,
                    uint64_t seed, uint32_t year_index, const double* cdf, const int32_t* cdf_idx,
                    const int32_t* root_slot, const int32_t* S_buf, const uint64_t* Q_buf,
                    const int32_t* q_idx, const int32_t* ns_buf, int32_t* cit, int8_t* kind,
                    int32_t* cnt, int32_t* near, int32_t* anext);

void launch_nb_draw_batch(hipStream_t s, const Batch& b, int m, int cap, uint32_t year_index,
                          const double* cdf, const int32_t* root_slot, const int32_t* S_buf,
                          const uint64_t* Q_buf, const int32_t* q_idx, const int32_t* ns_buf,
                          int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* near,
                          int32_t* anext);

// Stand-alone 2-hop neighbourhoods (out[i*cap ...], out_len[i]) by the
// workgroup path; ticks (nullable) gets per root the wall-clock ticks to the
// end of the 1-hop level, to the end, and the 2-hop candidate entries.
void launch_bfs_only(hipStream_t s, const CsrView& g, int64_t count, const int32_t* roots, int cap,
                     int32_t* out, int32_t* out_len, int64_t* ticks, const BfsStage& stage);

// ---- CSR append ----
// Sorted forward rows of the new papers appended at fwd_adj + fwd_base;
// src_of[e] = citing id; add[t] += 1 per edge.  new_off: int64 exclusive
// offsets of the per-paper counts (n_new+1).
void launch_pa_compact(hipStream_t s, int64_t n_exist, int64_t n_new, int m, const int32_t* cit,
                       const int32_t* cnt, const int64_t* new_off, int64_t fwd_base,
                       int32_t* fwd_off, int32_t* fwd_adj, int32_t* src_of);

// Stable LSD radix sort of (key,val) pairs by key (keys < 2^key_bits).
// Result lands in (k_out, v_out).  Buffers k_tmp/v_tmp same size as n.
int64_t radix_scratch_elems(int64_t n);
void launch_radix_sort_pairs(hipStream_t s, const uint32_t* k_in, const uint32_t* v_in,
                             uint32_t* k_out, uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp,
                             int64_t n, int key_bits, int64_t* scratch);

// Backward offsets and degrees after the year's citations, from their targets
// sorted ascending (keys[0, n_keys); padding keys >= n_old + n_new allowed):
// bwd_off_new[v] = bwd_off_old[min(v,n_old)] + #keys < v; in_deg[v] += #keys == v.
void launch_bwd_offsets_sorted(hipStream_t s, int64_t n_old, int64_t n_new,
                               const int32_t* bwd_off_old, const uint32_t* sorted_keys,
                               int64_t n_keys, int32_t* bwd_off_new, int32_t* in_deg);

// bwd_adj_new = old rows followed by the row's new citers (sorted_src,
// grouped by target in ascending citing order).  e_new < 0: the edge total
// is read on the device (bwd_off_new[n_total]) and e_bound sizes the grid.
// part (nullable, part_cap int64): scratch for the blocks' row ranges.
// merge_out: output positions per block, 256, 512, 1024 or 2048 (0: 1024).
void launch_bwd_merge(hipStream_t s, int64_t n_old, int64_t n_total, const int32_t* bwd_off_old,
                      const int32_t* bwd_adj_old, const int32_t* bwd_off_new,
                      const uint32_t* sorted_src, int32_t* bwd_adj_new, int64_t e_new,
                      int64_t e_bound, int64_t* part, int64_t part_cap, int merge_out = 0);

// Batched forward append (rows at the end of each simulation's region; the
// year's (target, citing) pairs also listed contiguously in stage/src_of).
void launch_pa_compact_batch(hipStream_t s, const Batch& b, int m, const int32_t* cit,
                             const int32_t* cnt, const int64_t* new_off, int32_t* fwd_off,
                             int32_t* fwd_adj, int32_t* src_of, int32_t* stage);
void launch_birth_batch(hipStream_t s, const Batch& b, int32_t* birth);
// keys[e] = v for e in [*from, to) (from read on the device).
void launch_fill_tail_u32(hipStream_t s, uint32_t* keys, const int64_t* from, int64_t to,
                          uint32_t v);

// ---- ER ----
void launch_er_count(hipStream_t s, int64_t n_exist, int64_t n_new, double p, uint64_t seed,
                     uint32_t year_index, int32_t* counts);
void launch_er_write(hipStream_t s, int64_t n_exist, int64_t n_new, double p, uint64_t seed,
                     uint32_t year_index, const int64_t* new_off, int64_t fwd_base,
                     int32_t* fwd_off, int32_t* fwd_adj, int32_t* src_of);

// Small helpers.
void launch_fill_i32(hipStream_t s, int32_t* p, int32_t v, int64_t n);
void launch_sum_i32(hipStream_t s, const int32_t* p, int64_t n, unsigned long long* out);
// *dst += *src (one thread).
void launch_acc_i64(hipStream_t s, const int64_t* src, unsigned long long* dst);

// ---- seed graphs generated on the device for an ensemble (k_gen.hip) ----
// One row per simulation: its first global id, its first flat paper index
// (papers of all simulations numbered in order), its seed papers, its first
// flat log-fitness index (its final papers), its edge region and its key.
struct GenSim {
    int64_t base, pfx, n0, lfx, ebase;
    uint64_t seed;
};
void launch_gen_hist(hipStream_t s, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                     int32_t* hist);
void launch_gen_counts(hipStream_t s, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                       int32_t year_lo, const int64_t* per_year, double lambda, double e0,
                       int32_t* birth, int32_t* cnt);
void launch_gen_sim_edges(hipStream_t s, const GenSim* t, int n_sims, const int64_t* off,
                          int64_t* sim_edges);
void launch_gen_rows(hipStream_t s, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                     int32_t year_lo, const int64_t* per_year, const int32_t* birth,
                     const int32_t* cnt, const int64_t* off, const int64_t* sim_edges,
                     int32_t* fwd_off, int32_t* fwd_adj, uint32_t* keys, uint32_t* vals);
void launch_gen_pad(hipStream_t s, const GenSim* t, int n_sims, int64_t tot,
                    const int64_t* sim_edges, int32_t* birth, int32_t* fwd_off);
void launch_gen_indeg(hipStream_t s, int64_t tot, const int32_t* bwd_off, int32_t* in_deg);
void launch_gen_logfit(hipStream_t s, const GenSim* t, int n_sims, int64_t n_flat, double mu,
                       double sigma, float* logfit);

// ---- PA likelihood (k_lik.hip) ----
constexpr int kLikMaxYears = 256;   // years per sequence (LDS per-year partials)
constexpr int kLikYearCols = 5;     // per year: Z, A, log L, alpha terms, tau terms
constexpr int kLikPPT = 4;          // papers per thread in the paper passes
constexpr int kLikSegPer = 8;       // sorted slots per thread in k_lik_segsum
constexpr int kLikSegChunk = 256 * kLikSegPer;

// A citation sequence in device memory.  Year y: papers [0, n_exist[y])
// exist, with in-degrees deg[deg_off[y] + j]; citing papers new_off[y] ..
// new_off[y] + n_new[y] - 1 (sequence-wide index p) fill cit/kind[p*m + s]
// for s < cnt[p].  Tables n_exist .. year_of are device arrays.
struct LikSeq {
    int32_t n_years;
    int32_t m;
    int64_t n_papers;
    const int64_t* n_exist;
    const int64_t* n_new;
    const int64_t* deg_off;
    const int64_t* new_off;
    const int32_t* year_of;
    const int32_t* deg;
    const int32_t* cit;
    const int8_t* kind;
    const int32_t* cnt;
    const int32_t* birth;
    int64_t n_cite;           // citing papers (sum of n_new)
    const int32_t* cit_t;     // slot-major copies: slot s of paper p at s*n_cite + p
    const int8_t* kind_t;
    const uint32_t* pos_t;    // by-target sorted position of each slot
    const float2* la_t;       // {log(k+1), age} of each slot's target (fixed by the sequence)
    const float* Lpy;         // log(k+1) of every paper-year (fixed by the sequence)
};

struct LikArgs {
    double alpha, tau;        // used when at_dev is null
    const double* at_dev;     // optional device {alpha, tau}
    const float* logfit;      // device [n_papers]
};

// Per-target sums of the slot shares: seg_off[n_papers+1] over the slots
// sorted by target, evt[n_papers] (runs inside one chunk), carry[2*chunks].
struct LikSegs {
    const int32_t* seg_off;
    double* evt;
    double* carry;
};

// Slot-pass shape for m slots per citing paper: G lanes per paper (a power
// of two, G*8 >= m), K = max(4, ceil(m/G)) slots per lane; a block of 256
// threads holds 256/G citing papers.
struct LikCiteShape {
    int G, K;
};
LikCiteShape lik_cite_shape(int m, int lanes = 0);
// la_t[s] for every slot s of the slot-major layout (zero for empty slots).
void launch_lik_slot_la(hipStream_t st, const LikSeq& q, float2* la_t);
// Lpy[i] = log(deg[i] + 1) for the n paper-years.
void launch_lik_deg_log(hipStream_t st, const int32_t* deg, int64_t n, float* Lpy);
// Whether `lanes` lanes per citing paper can hold m slots (lanes 1,2,4,8).
bool lik_lanes_ok(int m, int lanes);
int64_t lik_paper_blocks(int64_t n_papers);
int64_t lik_seg_chunks(int64_t S);
// Once per sequence: the T = new_papers*m slots sorted by target (empty
// slots last): skey = sorted keys, seg_off per paper, and the slot-major
// copies cit_t/kind_t/pos_t.  keys/vals/perm/k_tmp/v_tmp: T each; scratch:
// radix_scratch_elems(T).
void launch_lik_sort_slots(hipStream_t s, int64_t T, int m, int64_t n_papers,
                           const int32_t* cnt, const int32_t* cit, const int8_t* kind,
                           uint32_t* keys, uint32_t* vals, uint32_t* skey, uint32_t* perm,
                           int32_t* cit_t, int8_t* kind_t, uint32_t* pos_t, uint32_t* k_tmp,
                           uint32_t* v_tmp, int key_bits, int64_t* scratch, int32_t* seg_off);
// Paper pass (per-block partial Z_y), then Z_y.
void launch_lik_forward_sums(hipStream_t s, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                             double* part_p, double* year_sums);
// Slot pass (log L; with grad also A_y, the alpha/tau slot terms and each
// slot's dL/dphi share into share[pos_t[slot]]), then the per-year sums.
void launch_lik_slots(hipStream_t s, bool grad, LikCiteShape sh, int64_t nblk_e,
                      const int32_t* ev_block, const int32_t* year_blk, const LikSeq& q,
                      const LikArgs& a, double* year_sums, double* share, double* part_e);
// Segment sums of the shares over the S filled sorted slots, then the paper
// pass with A_y: dL/dphi = shares - sum_y A_y w_j(y); A-weighted partials.
void launch_lik_backward(hipStream_t s, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                         const double* year_sums, int64_t S, const uint32_t* skey,
                         const double* share, const LikSegs& sg, double* grad_phi,
                         double* part_p);
// {log L, dL/dalpha, dL/dtau} -> out[3] (and out2[3] if non-null).
void launch_lik_final(hipStream_t s, int n_years, bool grad, int64_t nblk_p,
                      const double* part_p, const LikArgs& a, const double* year_sums,
                      double* out, double* out2);

// ---- PA likelihood at B parameter points per pass (k_likb.hip) ----
constexpr int kLikBPPT = 8;      // papers per thread in the batched paper passes
constexpr int kLikBPaperBlocks = 2048;  // batched paper passes: blocks (each takes many tiles)
constexpr int kLikBCitePapers = 64;     // batched slot pass: citing papers per block
constexpr int kLikBSegPer = 16;  // sorted slots per lane group in the batched segment sums

// Buffers of one batched evaluation; per-point arrays interleave the B
// points (point b of item i at i*B + b).
struct LikBatchBufs {
    const double* at;          // {alpha_b, tau_b} [2B]
    const float* logfit;       // phi [n_papers * B]
    double* year_sums;         // [n_years * kLikYearCols * B]
    double* part;              // (n_years + 2) * nblk_p * B + 4 * nblk_e * B
    const int32_t* ev_block;   // slot-pass block table for the batched shape
    const int32_t* year_blk;
    int64_t nblk_p, nblk_e;
    double* share;             // [S * B] per-slot shares, by-target order
    int64_t S;                 // filled slots
    const uint32_t* skey;
    const int32_t* seg_off;
    double* evt;               // [n_papers * B] per-paper share totals
    // segment-sum jobs over the sorted positions, cut at paper boundaries:
    // job k covers [job_start[k], job_start[k+1]); job_tag[k] < 0 for a run
    // of whole papers (at most lik_batch_small_job(B) positions), else the
    // index of a piece (at most kLikBBigPiece positions) of one larger paper
    const int64_t* job_start;
    const int32_t* job_tag;
    int64_t n_jobs;
    double* bigpart;           // [pieces * B]
    const int32_t* big_paper;  // [n_big] the larger papers
    const int32_t* big_first;  // [n_big + 1] their first piece
    int32_t n_big;
    double* grad;              // dL/dphi [n_papers * B]
    double* out;               // {log L, dL/dalpha, dL/dtau} of point b at 3b [3B]
    double* out2;              // optional copy of out
};
bool lik_batch_ok(int B);                                 // B in {4, 8, 16}
bool lik_batch_shape_ok(int m, int B);                    // m slots fit the batched slot pass
LikCiteShape lik_batch_cite_shape(int m, int B);         // G lanes per (paper, point), K slots
int64_t lik_batch_paper_blocks(int64_t n_papers, int B);
int lik_batch_small_job(int B);                            // positions of a whole-paper job
constexpr int kLikBBigPiece = 4096;                        // positions of a larger paper's piece
// The whole evaluation (both paper passes, slot pass, segment sums, final);
// fast: hardware exp/log/reciprocal instead of the simulator's arithmetic.
void launch_lik_batch(hipStream_t s, int B, const LikSeq& q, const LikBatchBufs& bb,
                      LikCiteShape sh, bool fast);

}  // namespace pg
