// k_pa.hip -- PA citation draws without duplicates (phase 1 of a new paper).
//
// Replaces the reference's candidate sampling (MakePopulateCitations,
// src/abm.cu:461-565, run from Stage 2/3, src/kernel.cu:170-344), which walked
// a per-paper candidate bitmap with Efraimidis-Spirakis keys and a top-K
// heap, with this model's rule: draw u, take the first paper whose FP64 CDF
// value exceeds u*Z (binary search), and redraw on duplicates.
//
// Attempt a of paper g uses Philox counter (g, a, year index, TAG_PA), so the
// raw pick of every attempt is independent of the others.  A group of G
// lanes evaluates G consecutive attempts at once; attempt a is accepted iff
// no earlier attempt of the same paper produced the same paper id, which is
// exactly the sequential redraw-on-duplicate rule.  The group stops at the
// attempt that completes n_pa distinct picks and records where the next
// attempt starts (the refill in the neighbourhood kernel continues there).
#include "pg_internal.h"

namespace pg {
namespace {

// BATCH: paper i of the year's batch order belongs to the simulation found in
// the batch table, which supplies n_exist, the CDF region, the Philox key
// and n_pa; picks are written as global ids (region base + local id).
template <int G, bool BATCH>
__global__ __launch_bounds__(256) void k_pa_draw(Batch bt, int64_t n_exist,
                                                 const double* __restrict__ cdf,
                                                 const int32_t* __restrict__ cidx, int64_t n_new,
                                                 int m, int n_pa, uint64_t seed, uint32_t yi,
                                                 int a_max, int32_t* __restrict__ cit,
                                                 int8_t* __restrict__ kind, int32_t* __restrict__ cnt,
                                                 int32_t* __restrict__ near,
                                                 int32_t* __restrict__ anext) {
    constexpr int kGroups = 256 / G;
    __shared__ int32_t s_pick[256];
    __shared__ int32_t s_acc[kGroups][kMaxCites];
    const int grp = threadIdx.x / G;
    const int gl = threadIdx.x % G;
    const int lane = threadIdx.x & 63;
    const int shift = (lane / G) * G;  // first lane of this group in the wavefront
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << shift);
    const int64_t i = (int64_t)blockIdx.x * kGroups + grp;
    int64_t local = i;
    int64_t base = 0;
    if (BATCH) {
        n_new = bt.n_new;
        const BatchSim& sm = bt.sims[batch_sim_of_new(bt, i < n_new ? i : n_new - 1)];
        n_exist = sm.n_exist;
        base = sm.base;
        cdf += base;
        seed = sm.seed;
        n_pa = sm.n_pa;
        local = i - sm.new_first;
    }
    const bool live = i < n_new;
    const uint32_t gid = (uint32_t)(n_exist + local);
    const double Z = cdf[n_exist - 1];
    const double eps = kNearRel * Z;
    const double kscale = cdf_key_scale(Z);
    // no positive weight (every FP32 weight underflowed to zero): no PA
    // attempts, so the paper gets no citations (the neighbourhood draws and
    // the refill need a first pick)
    if (!(Z > 0.0)) a_max = 0;

    int c = 0, a0 = 0, nr = 0, a_stop = a_max;
    bool need = live && n_pa > 0 && a_max > 0;
    // a paper's G lanes and its LDS rows lie in one wavefront (G <= 64), so
    // the rounds synchronise wavefronts: a wavefront stops with its own
    // papers, not with the slowest paper of the block
    while (__ballot(need) != 0) {
        const int a = a0 + gl;
        const bool valid = need && a < a_max;
        int32_t pick = -1;
        bool nf = false;
        if (valid) {
            u32x4 r = philox(gid, (uint32_t)a, yi, TAG_PA, seed);
            double t = u53(r.x, r.y) * Z;
            int64_t k = cdf_search(cdf, cidx, n_exist, t, kscale);
            nf = cdf_near(cdf, n_exist, k, t, eps);
            pick = (int32_t)k;
        }
        s_pick[threadIdx.x] = pick;
        wave_lds_sync();
        bool dup = !valid;
        if (valid) {
            const int g0 = grp * G;
            for (int b = 0; b < gl && !dup; ++b) dup = (s_pick[g0 + b] == pick);
            for (int b = 0; b < c && !dup; ++b) dup = (s_acc[grp][b] == pick);
        }
        const uint64_t acc_mask = __ballot(!dup) & gmask;
        const uint64_t near_mask = __ballot(nf) & gmask;
        const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        const int pos = c + __popcll(acc_mask & below);
        const bool take = !dup && pos < n_pa;
        if (take) {
            cit[i * m + pos] = (int32_t)(pick + base);
            kind[i * m + pos] = 0;
            s_acc[grp][pos] = pick;
        }
        if (need) {
            const int got = __popcll(acc_mask);
            int used = G;  // attempts consumed in this round
            if (c + got >= n_pa) {
                // lane holding the n_pa-th acceptance
                uint64_t mm = acc_mask >> shift;
                int left = n_pa - c;
                int l = 0;
                for (; l < G; ++l) {
                    if ((mm >> l) & 1ull) { if (--left == 0) break; }
                }
                used = l + 1;
                a_stop = a0 + used;
            }
            const uint64_t used_mask = (used >= 64) ? ~0ull : (((1ull << used) - 1ull) << shift);
            nr += __popcll(near_mask & used_mask);
            c = (c + got < n_pa) ? c + got : n_pa;
            a0 += G;
            if (c >= n_pa || a0 >= a_max) {
                need = false;
                if (c < n_pa) a_stop = a_max;
            }
        }
        wave_lds_sync();
    }
    if (live && gl == 0) {
        for (int j = c; j < m; ++j) { cit[i * m + j] = -1; kind[i * m + j] = -1; }
        cnt[i] = c;
        near[i] = nr;
        anext[i] = (n_pa > 0) ? a_stop : 0;
    }
}

template <bool BATCH>
void pa_draw_g(hipStream_t s, int G, const Batch& bt, int64_t n_exist, const double* cdf,
               const int32_t* cidx, int64_t n_new, int m, int n_pa, uint64_t seed,
               uint32_t year_index, int32_t* cit,
               int8_t* kind, int32_t* cnt, int32_t* near, int32_t* anext) {
    const int a_max = kAttemptsPerCite * m;
    const int64_t per = 256 / G;
    const dim3 grid((uint32_t)((n_new + per - 1) / per)), block(256);
    switch (G) {
        case 8: hipLaunchKernelGGL((k_pa_draw<8, BATCH>), grid, block, 0, s, bt, n_exist, cdf, cidx, n_new,
                                   m, n_pa, seed, year_index, a_max, cit, kind, cnt, near, anext); break;
        case 16: hipLaunchKernelGGL((k_pa_draw<16, BATCH>), grid, block, 0, s, bt, n_exist, cdf, cidx, n_new,
                                    m, n_pa, seed, year_index, a_max, cit, kind, cnt, near, anext); break;
        case 32: hipLaunchKernelGGL((k_pa_draw<32, BATCH>), grid, block, 0, s, bt, n_exist, cdf, cidx, n_new,
                                    m, n_pa, seed, year_index, a_max, cit, kind, cnt, near, anext); break;
        default: hipLaunchKernelGGL((k_pa_draw<64, BATCH>), grid, block, 0, s, bt, n_exist, cdf, cidx, n_new,
                                    m, n_pa, seed, year_index, a_max, cit, kind, cnt, near, anext); break;
    }
}

}  // namespace

int pa_draw_lanes(int n_pa) { return n_pa <= 8 ? 8 : n_pa <= 16 ? 16 : n_pa <= 32 ? 32 : 64; }

void launch_pa_draw(hipStream_t s, int64_t n_exist, const double* cdf, const int32_t* cdf_idx,
                    int64_t n_new, int m, int n_pa, uint64_t seed, uint32_t year_index,
                    int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* near, int32_t* anext,
                    int lanes) {
    if (n_new <= 0) return;
    const int G = lanes > 0 ? lanes : pa_draw_lanes(n_pa);
    pa_draw_g<false>(s, G, Batch{}, n_exist, cdf, cdf_idx, n_new, m, n_pa, seed, year_index, cit,
                     kind, cnt, near, anext);
}

void launch_pa_draw_batch(hipStream_t s, const Batch& b, int m, uint32_t year_index, int lanes,
                          const double* cdf, int32_t* cit, int8_t* kind, int32_t* cnt,
                          int32_t* near, int32_t* anext) {
    if (b.n_new <= 0) return;
    pa_draw_g<true>(s, lanes, b, 0, cdf, nullptr, b.n_new, m, 0, 0, year_index, cit, kind, cnt,
                    near, anext);
}

}  // namespace pg
