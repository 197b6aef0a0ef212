// k_gen.hip -- synthetic seed graphs generated on the device for a batched
// ensemble: every simulation's graph and log-fitness, bit for bit what the
// host generator (pg_seed_graph, pg_seed_logfit) gives for the simulation's
// Philox key, written straight into the ensemble's arrays (global ids, each
// simulation in its id and edge regions).  The draws are the shared
// functions of pg_math.h; the backward CSR is the transpose built by the
// year step's radix sort and offsets kernels.
#include <hip/hip_runtime.h>

#include "pg_internal.h"
#include "pg_math.h"

namespace pg {

namespace {

constexpr int kGT = 256;

__device__ __forceinline__ int gen_sim_of(const GenSim* __restrict__ t, int n, int64_t x,
                                          bool by_logfit) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((by_logfit ? t[mid].lfx : t[mid].pfx) <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Histogram of raw birth years per simulation (integer atomics: exact).
__global__ __launch_bounds__(kGT) void k_gen_hist(const GenSim* __restrict__ t, int n_sims,
                                                  int64_t n_flat, int32_t span,
                                                  int32_t* __restrict__ hist) {
    const int64_t x = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (x >= n_flat) return;
    const int s = gen_sim_of(t, n_sims, x, false);
    const int64_t j = x - t[s].pfx;
    atomicAdd(&hist[(int64_t)s * span + seed_birth_offset(t[s].seed, (uint32_t)j, span)], 1);
}

// Births (ids sorted by year) and citation counts.  per_year[s*(span+1)+y]
// = papers of simulation s born before year_lo + y.
__global__ __launch_bounds__(kGT) void k_gen_counts(const GenSim* __restrict__ t, int n_sims,
                                                    int64_t n_flat, int32_t span, int32_t year_lo,
                                                    const int64_t* __restrict__ per_year,
                                                    double lambda, double e0,
                                                    int32_t* __restrict__ birth,
                                                    int32_t* __restrict__ cnt) {
    const int64_t x = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (x >= n_flat) return;
    const int s = gen_sim_of(t, n_sims, x, false);
    const int64_t i = x - t[s].pfx;
    const int64_t* py = per_year + (int64_t)s * (span + 1);
    int y = 0;
    while (y + 1 < span && py[y + 1] <= i) ++y;
    birth[t[s].base + i] = year_lo + y;
    cnt[x] = seed_count(t[s].seed, (uint32_t)i, lambda, e0, py[y]);
}

// Per simulation: its first flat edge offset and edge count.
__global__ void k_gen_sim_edges(const GenSim* __restrict__ t, int n_sims,
                                const int64_t* __restrict__ off, int64_t* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sims) return;
    const int64_t a = off[t[s].pfx], b = off[t[s].pfx + t[s].n0];
    out[2 * s] = a;
    out[2 * s + 1] = b - a;
}

// Forward rows: paper i's draws (distinct, draw order) into its row, then
// sorted ascending; its row start in fwd_off; the (target, citer) pairs at
// the flat edge positions for the transpose.
__global__ __launch_bounds__(kGT) void k_gen_rows(const GenSim* __restrict__ t, int n_sims,
                                                  int64_t n_flat, int32_t span, int32_t year_lo,
                                                  const int64_t* __restrict__ per_year,
                                                  const int32_t* __restrict__ birth,
                                                  const int32_t* __restrict__ cnt,
                                                  const int64_t* __restrict__ off,
                                                  const int64_t* __restrict__ sim_edges,
                                                  int32_t* __restrict__ fwd_off,
                                                  int32_t* __restrict__ fwd_adj,
                                                  uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
    const int64_t x = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (x >= n_flat) return;
    const int s = gen_sim_of(t, n_sims, x, false);
    const GenSim g = t[s];
    const int64_t i = x - g.pfx;
    const int64_t e_loc = off[x] - sim_edges[2 * s];
    const int64_t row = g.ebase + e_loc;
    fwd_off[g.base + i] = (int32_t)row;
    const int c = cnt[x];
    if (c == 0) return;
    const int64_t older = per_year[(int64_t)s * (span + 1) + (birth[g.base + i] - year_lo)];
    int32_t* out = fwd_adj + row;
    int got = 0;
    for (uint32_t a = 0; got < c; ++a) {
        const int32_t tg = seed_target(g.seed, (uint32_t)i, a, older);
        bool dup = false;
        for (int q = 0; q < got && !dup; ++q) dup = out[q] == tg;
        if (!dup) out[got++] = tg;
    }
    for (int q = 1; q < c; ++q) {  // insertion sort of the short row
        const int32_t v = out[q];
        int r = q - 1;
        while (r >= 0 && out[r] > v) {
            out[r + 1] = out[r];
            --r;
        }
        out[r + 1] = v;
    }
    for (int q = 0; q < c; ++q) {
        const int32_t gid = (int32_t)(g.base + out[q]);
        out[q] = gid;
        keys[off[x] + q] = (uint32_t)gid;
        vals[off[x] + q] = (uint32_t)(g.base + i);
    }
}

// Ids past a simulation's seed papers in its region: no birth yet, and
// forward offsets at the end of the seed rows (fwd_off[tot] for the last
// simulation's end).
__global__ __launch_bounds__(kGT) void k_gen_pad(const GenSim* __restrict__ t, int n_sims,
                                                 int64_t tot, const int64_t* __restrict__ sim_edges,
                                                 int32_t* __restrict__ birth,
                                                 int32_t* __restrict__ fwd_off) {
    const int64_t v = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (v > tot) return;
    if (v == tot) {
        const GenSim g = t[n_sims - 1];
        fwd_off[tot] = (int32_t)(g.ebase + sim_edges[2 * (n_sims - 1) + 1]);
        return;
    }
    int lo = 0, hi = n_sims - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].base <= v) lo = mid; else hi = mid - 1;
    }
    const GenSim g = t[lo];
    if (v - g.base >= g.n0) {
        birth[v] = 0;
        fwd_off[v] = (int32_t)(g.ebase + sim_edges[2 * lo + 1]);
    }
}

__global__ __launch_bounds__(kGT) void k_gen_indeg(int64_t tot, const int32_t* __restrict__ bwd_off,
                                                   int32_t* __restrict__ in_deg) {
    const int64_t v = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (v < tot) in_deg[v] = bwd_off[v + 1] - bwd_off[v];
}

__global__ __launch_bounds__(kGT) void k_gen_logfit(const GenSim* __restrict__ t, int n_sims,
                                                    int64_t n_flat, double mu, double sigma,
                                                    float* __restrict__ logfit) {
    const int64_t x = (int64_t)blockIdx.x * kGT + threadIdx.x;
    if (x >= n_flat) return;
    const int s = gen_sim_of(t, n_sims, x, true);
    const int64_t k = x - t[s].lfx;
    logfit[t[s].base + k] = seed_logfit(t[s].seed, (uint32_t)k, mu, sigma);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kGT - 1) / kGT); }

}  // namespace

void launch_gen_hist(hipStream_t st, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                     int32_t* hist) {
    if (n_flat > 0) k_gen_hist<<<grid_of(n_flat), kGT, 0, st>>>(t, n_sims, n_flat, span, hist);
}

void launch_gen_counts(hipStream_t st, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                       int32_t year_lo, const int64_t* per_year, double lambda, double e0,
                       int32_t* birth, int32_t* cnt) {
    if (n_flat > 0)
        k_gen_counts<<<grid_of(n_flat), kGT, 0, st>>>(t, n_sims, n_flat, span, year_lo, per_year,
                                                     lambda, e0, birth, cnt);
}

void launch_gen_sim_edges(hipStream_t st, const GenSim* t, int n_sims, const int64_t* off,
                          int64_t* sim_edges) {
    k_gen_sim_edges<<<(n_sims + 63) / 64, 64, 0, st>>>(t, n_sims, off, sim_edges);
}

void launch_gen_rows(hipStream_t st, const GenSim* t, int n_sims, int64_t n_flat, int32_t span,
                     int32_t year_lo, const int64_t* per_year, const int32_t* birth,
                     const int32_t* cnt, const int64_t* off, const int64_t* sim_edges,
                     int32_t* fwd_off, int32_t* fwd_adj, uint32_t* keys, uint32_t* vals) {
    if (n_flat > 0)
        k_gen_rows<<<grid_of(n_flat), kGT, 0, st>>>(t, n_sims, n_flat, span, year_lo, per_year,
                                                   birth, cnt, off, sim_edges, fwd_off, fwd_adj,
                                                   keys, vals);
}

void launch_gen_pad(hipStream_t st, const GenSim* t, int n_sims, int64_t tot,
                    const int64_t* sim_edges, int32_t* birth, int32_t* fwd_off) {
    k_gen_pad<<<grid_of(tot + 1), kGT, 0, st>>>(t, n_sims, tot, sim_edges, birth, fwd_off);
}

void launch_gen_indeg(hipStream_t st, int64_t tot, const int32_t* bwd_off, int32_t* in_deg) {
    if (tot > 0) k_gen_indeg<<<grid_of(tot), kGT, 0, st>>>(tot, bwd_off, in_deg);
}

void launch_gen_logfit(hipStream_t st, const GenSim* t, int n_sims, int64_t n_flat, double mu,
                       double sigma, float* logfit) {
    if (n_flat > 0)
        k_gen_logfit<<<grid_of(n_flat), kGT, 0, st>>>(t, n_sims, n_flat, mu, sigma, logfit);
}

}  // namespace pg
