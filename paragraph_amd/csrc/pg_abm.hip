// pg_abm.hip -- host side of the reference's scoring model: configuration
// (pg_abm_configure), the year step behind pg_year_step(PG_MODEL_ABM), and
// the downloads its tests and the `abm` driver use.
//
// A year, all stream-ordered on the context's stream:
//   fitness of papers not yet drawn -> score sums and components (k_abm.hip)
//   -> same-year keys, radix sort, subset -> selection (paper x level)
//   -> row counts, scan -> (host: edge total, 2^21 flag) -> rows
//   -> the CSR append shared with the PA and ER years (radix sort by target,
//      backward offsets, backward merge).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_ctx.h"
#include "pg_internal.h"

using namespace pg;

namespace {

std::vector<double> fitness_table() {
    std::vector<double> c(kAbmFitMax);
    double s = 0.0;
    for (int f = 1; f <= kAbmFitMax; ++f) {
        s += 6.3742991333 * 0.072 * std::pow((double)f, -1.634);
        c[f - 1] = s;
    }
    return c;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

void abm_release(pg_ctx* c) {
    pg_abm_store& a = c->abm;
    dfree(a.bag); dfree(a.rec); dfree(a.cdf); dfree(a.fit); dfree(a.comp); dfree(a.parts);
    dfree(a.cnt); dfree(a.rt); dfree(a.tot); dfree(a.syt); dfree(a.sy_keys); dfree(a.sy_vals);
    dfree(a.sel); dfree(a.nsel); dfree(a.info); dfree(a.cand); dfree(a.kc);
    a.on = false;
}

int abm_year_step(pg_ctx* ctx, const pg_year_params* p, pg_year_stats* stats) {
    pg_abm_store& a = ctx->abm;
    if (!a.on) return fail(ctx, PG_ESTATE, "pg_year_step: PG_MODEL_ABM before pg_abm_configure");
    if (ctx->seq_recording)
        return fail(ctx, PG_EINVAL, "pg_year_step: the sequence store records PA years only");
    const int64_t N = ctx->n, D = p->n_new;
    if (N < 1) return fail(ctx, PG_EINVAL, "pg_year_step: the scoring model needs a paper to cite");
    if (N + D > (int64_t)INT32_MAX) return fail(ctx, PG_ECAP, "pg_year_step: ids exceed int32");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    bool used[PG_K_COUNT] = {}, used2[PG_K_COUNT] = {};
    auto on = [&](int k) { return ctx->prof && ((ctx->prof_mask >> k) & 1u); };
    auto ev = [&](int k, int j) -> hipError_t {
        if (!on(k)) return hipSuccess;
        if (j == 1) used[k] = true;
        return hipEventRecord(ctx->pev[k][j], s);
    };
    auto ev2 = [&](int k, int j) -> hipError_t {
        if (!on(k)) return hipSuccess;
        if (j == 1) used2[k] = true;
        return hipEventRecord(ctx->pev2[k][j], s);
    };
    PG_HIP(hipEventRecord(ctx->ev0, s));
    int64_t e_year = 0;
    if (D > 0) {
        if (int rc = ensure_year_edges(ctx, D)) return rc;
        PG_HIP(ev(PG_K_ABM_SCORES, 0));
        if (a.fit_n < N + D) {
            launch_abm_fit(s, a.fit_n, N + D, ctx->cfg.seed, a.cdf, a.fit);
            a.fit_n = N + D;
        }
        launch_fill_i32(s, ctx->birth + N, p->year, D);
        launch_fill_i32(s, a.cnt, 0, kAbmRecMax + 1);
        launch_fill_i32(s, reinterpret_cast<int32_t*>(a.cand), 0, 4);
        launch_abm_scores(s, N, ctx->in_deg, ctx->birth, a.fit, p->year, a.n_rec, a.rec, a.parts,
                          a.cnt, a.cnt + kAbmRecMax, a.tot, a.rt, a.comp);
        const uint32_t yi = (uint32_t)p->year_index;
        launch_abm_same_keys(s, N, D, ctx->cfg.seed, yi, a.sy_keys, a.sy_vals, a.syt);
        const int64_t K = (int64_t)std::floor((double)D * a.syp);
        if (D >= 2 && K > 0) {
            launch_radix_sort_pairs(s, a.sy_keys, a.sy_vals, ctx->k_sorted, ctx->v_sorted,
                                    ctx->k_tmp, ctx->v_tmp, D, 32, ctx->scratch);
            launch_abm_same_mark(s, N, D, K, ctx->cfg.seed, yi, ctx->v_sorted, a.syt);
        }
        PG_HIP(ev(PG_K_ABM_SCORES, 1));
        AbmYear y{ctx->cfg.seed, N, D, yi, p->year, a.bag, a.n_bag, a.alpha, a.w_pa, a.w_rec,
                  a.w_fit, a.frc};
        PG_HIP(ev(PG_K_ABM_SELECT, 0));
        launch_abm_select(s, y, view(ctx), a.comp, a.syt, a.sel, a.nsel, a.kc);
        launch_sum_i32(s, a.kc, D, a.cand);
        launch_sum_i32(s, a.kc + D, D, a.cand + 1);
        PG_HIP(ev(PG_K_ABM_SELECT, 1));
        PG_HIP(ev(PG_K_ABM_ROWS, 0));
        launch_abm_count(s, y, a.syt, a.nsel, ctx->cnt, a.info);
        launch_excl_scan_i32(s, ctx->cnt, D, ctx->new_off, ctx->scratch);
        PG_HIP(ev(PG_K_ABM_ROWS, 1));
        PG_HIP(hipGetLastError());
        PG_HIP(hipMemcpyAsync(ctx->h_sum, ctx->new_off + D, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        PG_HIP(hipMemcpyAsync(ctx->h_sum + 1, a.cnt + kAbmRecMax, sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
        PG_HIP(hipMemcpyAsync(ctx->h_sum + 2, a.cand, 2 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s));
        PG_HIP(hipStreamSynchronize(s));
        a.last_cand[0] = (int64_t)ctx->h_sum[2];
        a.last_cand[1] = (int64_t)ctx->h_sum[3];
        e_year = (int64_t)ctx->h_sum[0];
        if ((int32_t)ctx->h_sum[1] != 0)
            return fail(ctx, PG_ECAP, "pg_year_step: an in-degree reached 2^21 (k^3 overflows)");
        if (ctx->e + e_year > ctx->cfg.max_edges)
            return fail(ctx, PG_ECAP, "pg_year_step: edges would exceed max_edges");
        if (ctx->e + e_year > (int64_t)INT32_MAX)
            return fail(ctx, PG_ECAP, "pg_year_step: edge offsets exceed int32");
        if (int rc = ensure_year_edges(ctx, std::max(e_year, D))) return rc;
        PG_HIP(ev2(PG_K_ABM_ROWS, 0));
        launch_abm_write(s, y, a.syt, a.sel, a.info, ctx->new_off, ctx->e, ctx->fwd_off,
                         ctx->fwd_adj, ctx->src_of);
        PG_HIP(ev2(PG_K_ABM_ROWS, 1));
        PG_HIP(ev(PG_K_CSR_APPEND, 0));
        uint32_t* keys = reinterpret_cast<uint32_t*>(ctx->fwd_adj + ctx->e);
        launch_radix_sort_pairs(s, keys, reinterpret_cast<const uint32_t*>(ctx->src_of),
                                ctx->k_sorted, ctx->v_sorted, ctx->k_tmp, ctx->v_tmp, e_year,
                                key_bits_for(N + D), ctx->scratch);
        const int nxt = 1 - ctx->cur;
        launch_bwd_offsets_sorted(s, N, D, ctx->bwd_off[ctx->cur], ctx->k_sorted, e_year,
                                  ctx->bwd_off[nxt], ctx->in_deg);
        launch_bwd_merge(s, N, N + D, ctx->bwd_off[ctx->cur], ctx->bwd_adj[ctx->cur],
                         ctx->bwd_off[nxt], ctx->v_sorted, ctx->bwd_adj[nxt], ctx->e + e_year, 0,
                         ctx->scratch, ctx->cfg.max_papers + 2, ctx->tune_merge_out);
        PG_HIP(ev(PG_K_CSR_APPEND, 1));
        PG_HIP(hipGetLastError());
        ctx->cur = nxt;
    }
    PG_HIP(hipEventRecord(ctx->ev1, s));
    PG_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    PG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ctx->prof && D > 0) {
        // algorithmic bytes: scores read in_deg, birth (4 B) and fitness (2 B)
        // in each of two passes and write P,R,F (16 B); the same-year keys are
        // 8 B per paper plus four radix passes of 16 B; the selection reads
        // each 1-hop entry twice (4 B, once per level), gathers its P,R,F
        // (16 B) and its row offsets (16 B), and each 2-hop entry (4 B) with
        // its P,R,F (16 B); rows read the counts and picks and write 8 B per
        // edge (target, source for the append)
        const double e_old = (double)ctx->e;
        double bytes[PG_K_COUNT] = {};
        bytes[PG_K_ABM_SCORES] = 36.0 * N + 72.0 * D;
        bytes[PG_K_ABM_SELECT] = 40.0 * (double)a.last_cand[0] + 20.0 * (double)a.last_cand[1];
        bytes[PG_K_ABM_ROWS] = 12.0 * e_year + 80.0 * D;
        bytes[PG_K_CSR_APPEND] = 4.0 * (e_year + D) + 4.0 * (e_old + N) +
                                 4.0 * (e_old + e_year + N + D) + 8.0 * e_year;
        for (int k = 0; k < PG_K_COUNT; ++k) {
            if (!used[k]) continue;
            float kms = 0.0f;
            PG_HIP(hipEventElapsedTime(&kms, ctx->pev[k][0], ctx->pev[k][1]));
            if (used2[k]) {
                float k2 = 0.0f;
                PG_HIP(hipEventElapsedTime(&k2, ctx->pev2[k][0], ctx->pev2[k][1]));
                kms += k2;
            }
            ctx->prof_ms[k] += kms;
            ctx->prof_bytes[k] += bytes[k];
            ctx->prof_calls[k] += 1;
        }
    }
    if (stats) {
        stats->n_before = N;
        stats->n_after = N + D;
        stats->edges_added = e_year;
        stats->near_boundary = 0;
        stats->shortfall = 0;
        stats->pa_attempts = 0;
        stats->device_ms = ms;
    }
    a.last_n = N;
    a.last_d = D;
    ctx->n = N + D;
    ctx->e += e_year;
    ctx->last_n_new = D;
    ctx->last_n_exist = N;
    ctx->last_m = 0;
    ctx->last_model = PG_MODEL_ABM;
    ctx->steps_done += 1;
    ctx->last_year = p->year;
    ctx->last_year_index = p->year_index;
    return PG_OK;
}

extern "C" {

void pg_abm_fitness_table(double* cdf) {
    const std::vector<double> c = fitness_table();
    std::memcpy(cdf, c.data(), c.size() * sizeof(double));
}

int pg_abm_configure(pg_ctx* ctx, const pg_abm_model* m) {
    if (!ctx) return PG_EINVAL;
    if (!m || m->size != sizeof(pg_abm_model))
        return fail(ctx, PG_EINVAL, "pg_abm_configure: null model or struct size mismatch");
    if (!m->bag || m->n_bag < 1) return fail(ctx, PG_EINVAL, "pg_abm_configure: empty out-degree bag");
    for (int32_t i = 0; i < m->n_bag; ++i)
        if (m->bag[i] < 0 || m->bag[i] > kAbmMaxOut)
            return fail(ctx, PG_EINVAL, "pg_abm_configure: out-degree bag values must be in [0, 250]");
    if (m->n_rec < 0 || m->n_rec > kAbmRecMax || (m->n_rec > 0 && !m->rec_prob))
        return fail(ctx, PG_EINVAL, "pg_abm_configure: recency table must have 0..1024 entries");
    for (int32_t d = 0; d < m->n_rec; ++d)
        if (!(m->rec_prob[d] >= 0.0) || std::isinf(m->rec_prob[d]))
            return fail(ctx, PG_EINVAL, "pg_abm_configure: recency probabilities must be finite and >= 0");
    if (!(m->alpha <= 1.0) || std::isnan(m->alpha))
        return fail(ctx, PG_EINVAL, "pg_abm_configure: alpha must be <= 1 (negative: per-paper uniform)");
    const bool wrand = m->w_pa == -1.0 || m->w_rec == -1.0 || m->w_fit == -1.0;
    if (!wrand && !(m->w_pa >= 0.0 && m->w_rec >= 0.0 && m->w_fit >= 0.0 &&
                    m->w_pa + m->w_rec + m->w_fit > 0.0 && std::isfinite(m->w_pa + m->w_rec + m->w_fit)))
        return fail(ctx, PG_EINVAL, "pg_abm_configure: weights must be >= 0 with a positive sum, or -1");
    if (!(m->fully_random >= 0.0 && m->fully_random <= 1.0) ||
        !(m->same_year >= 0.0 && m->same_year <= 1.0))
        return fail(ctx, PG_EINVAL, "pg_abm_configure: fully_random and same_year must be in [0,1]");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    abm_release(ctx);
    pg_abm_store& a = ctx->abm;
    const int64_t NP = ctx->cfg.max_papers, DN = std::max<int64_t>(ctx->cfg.max_new_per_year, 1);
    hipError_t r = dalloc(&a.bag, m->n_bag);
    if (r == hipSuccess) r = dalloc(&a.rec, kAbmRecMax);
    if (r == hipSuccess) r = dalloc(&a.cdf, kAbmFitMax);
    if (r == hipSuccess) r = dalloc(&a.fit, NP);
    if (r == hipSuccess) r = dalloc(&a.comp, NP);
    if (r == hipSuccess) r = dalloc(&a.parts, 3 * kAbmMaxParts);
    if (r == hipSuccess) r = dalloc(&a.cnt, kAbmRecMax + 1);
    if (r == hipSuccess) r = dalloc(&a.rt, kAbmRecMax);
    if (r == hipSuccess) r = dalloc(&a.tot, 2);
    if (r == hipSuccess) r = dalloc(&a.syt, DN);
    if (r == hipSuccess) r = dalloc(&a.sy_keys, DN);
    if (r == hipSuccess) r = dalloc(&a.sy_vals, DN);
    if (r == hipSuccess) r = dalloc(&a.sel, 2 * DN * kAbmSlots);
    if (r == hipSuccess) r = dalloc(&a.nsel, 2 * DN);
    if (r == hipSuccess) r = dalloc(&a.info, 6 * DN);
    if (r == hipSuccess) r = dalloc(&a.cand, 2);
    if (r == hipSuccess) r = dalloc(&a.kc, 2 * DN);
    if (r != hipSuccess) {
        abm_release(ctx);
        return fail(ctx, PG_ENOMEM, "pg_abm_configure: device allocation failed");
    }
    a.h_bag.assign(m->bag, m->bag + m->n_bag);
    a.h_cdf = fitness_table();
    std::vector<double> rec(kAbmRecMax, 0.0);
    for (int32_t d = 0; d < m->n_rec; ++d) rec[d] = m->rec_prob[d];
    PG_HIP(hipMemcpy(a.bag, m->bag, (size_t)m->n_bag * sizeof(int32_t), hipMemcpyHostToDevice));
    PG_HIP(hipMemcpy(a.rec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
    PG_HIP(hipMemcpy(a.cdf, a.h_cdf.data(), a.h_cdf.size() * sizeof(double), hipMemcpyHostToDevice));
    a.n_bag = m->n_bag;
    a.n_rec = m->n_rec;
    a.alpha = m->alpha;
    a.w_pa = m->w_pa;
    a.w_rec = m->w_rec;
    a.w_fit = m->w_fit;
    a.frc = m->fully_random;
    a.syp = m->same_year;
    a.fit_n = 0;
    a.last_n = -1;
    a.last_d = 0;
    a.on = true;
    return PG_OK;
}

int pg_abm_year_info(pg_ctx* ctx, int32_t* info) {
    if (!ctx) return PG_EINVAL;
    const pg_abm_store& a = ctx->abm;
    if (!a.on || a.last_n < 0 || ctx->last_model != PG_MODEL_ABM)
        return fail(ctx, PG_ESTATE, "pg_abm_year_info: the last year was not a PG_MODEL_ABM year");
    if (!info) return fail(ctx, PG_EINVAL, "pg_abm_year_info: null buffer");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (a.last_d > 0)
        PG_HIP(hipMemcpy(info, a.info, (size_t)a.last_d * 6 * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_abm_year_counts(pg_ctx* ctx, int64_t* counts) {
    if (!ctx) return PG_EINVAL;
    const pg_abm_store& a = ctx->abm;
    if (!a.on || a.last_n < 0 || ctx->last_model != PG_MODEL_ABM)
        return fail(ctx, PG_ESTATE, "pg_abm_year_counts: the last year was not a PG_MODEL_ABM year");
    if (!counts) return fail(ctx, PG_EINVAL, "pg_abm_year_counts: null buffer");
    counts[0] = a.last_cand[0];
    counts[1] = a.last_cand[1];
    return PG_OK;
}

int pg_abm_scores(pg_ctx* ctx, float* comp) {
    if (!ctx) return PG_EINVAL;
    const pg_abm_store& a = ctx->abm;
    if (!a.on || a.last_n < 0 || ctx->last_model != PG_MODEL_ABM)
        return fail(ctx, PG_ESTATE, "pg_abm_scores: the last year was not a PG_MODEL_ABM year");
    if (!comp) return fail(ctx, PG_EINVAL, "pg_abm_scores: null buffer");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (a.last_d > 0)
        PG_HIP(hipMemcpy(comp, a.comp, (size_t)a.last_n * sizeof(float4), hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_abm_paper_draws(pg_ctx* ctx, int64_t first, int64_t count, int32_t* out_degree,
                       double* alpha, double* weights, int32_t* fitness) {
    if (!ctx) return PG_EINVAL;
    const pg_abm_store& a = ctx->abm;
    if (!a.on) return fail(ctx, PG_ESTATE, "pg_abm_paper_draws: no model configured");
    if (first < 0 || count < 0 || first + count > (int64_t)UINT32_MAX)
        return fail(ctx, PG_EINVAL, "pg_abm_paper_draws: bad id range");
    const uint64_t seed = ctx->cfg.seed;
    for (int64_t i = 0; i < count; ++i) {
        const uint32_t id = (uint32_t)(first + i);
        if (out_degree || alpha || weights) {
            const AbmPaper p = abm_paper(seed, id, a.h_bag.data(), a.n_bag, a.alpha, a.w_pa,
                                         a.w_rec, a.w_fit);
            if (out_degree) out_degree[i] = p.k;
            if (alpha) alpha[i] = p.alpha;
            if (weights)
                for (int j = 0; j < 3; ++j) weights[3 * i + j] = p.w[j];
        }
        if (fitness) fitness[i] = abm_fitness(seed, id, a.h_cdf.data());
    }
    return PG_OK;
}

}  // extern "C"
