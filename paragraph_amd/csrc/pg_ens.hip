// pg_ens.hip -- batched ensembles: many independent PA simulations advanced
// together (C ABI in include/paragraph.h; no reference counterpart -- the
// reference runs one simulation per process).
//
// All simulations share one global paper-id space.  Simulation s owns ids
// [R_s, R_s + RS_s) with RS_s = its final paper count + 1 rounded up to a
// multiple of kTile: its CDF tiles fall exactly where a plain run's do (so
// its FP64 CDF bits are a plain run's), and the spare id after its last
// paper keeps fwd_off[last + 1] equal to that paper's row end.  Its forward
// rows live in its own edge region [EB_s, EB_s + EC_s) (seed edges + all
// future citations), appended at the region's current end; the backward CSR
// is one global array in id order, rebuilt by the ordinary merge each year.
// Philox counters use local ids, so simulation s reproduces a plain run with
// the same seed graph, parameters and key bit for bit.
//
// One year step is a fixed sequence of launches over all simulations (sizes
// from the deterministic growth schedule; data-dependent counts stay on the
// device), so a whole run is captured once into a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_ctx.h"
#include "pg_internal.h"

using namespace pg;

struct pg_ensemble {
    pg_ensemble_config cfg{};
    std::vector<pg_sim_params> sims;
    pg_ctx* ctx = nullptr;                // arrays over the global id space
    std::string err;
    int64_t TOT = 0, ETOT = 0;
    std::vector<int64_t> R, RS, EB, EC, nfinal;
    std::vector<std::vector<int64_t>> D;   // D[s][y]
    std::vector<int64_t> year_new, year_tiles;
    std::vector<BatchSim> host_tab;        // [years][sims]
    BatchSim* dev_tab = nullptr;
    int32_t* stage = nullptr;              // the year's (target) list for the radix sort
    unsigned long long* counters = nullptr;  // near, attempts, BFS bytes, big count | edges
    int loaded_upto = 0;                   // simulations loaded, in order
    int64_t bwd_fill = 0;                  // backward entries loaded so far
    bool ready = false;                    // all loaded and snapshotted
    bool fresh = false;                    // state is the seed state
    int next_year = 0;                     // growth years done since the seed state
    int pa_lanes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

std::string g_ens_err;

int efail(pg_ensemble* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    else g_ens_err = msg;
    return code;
}

#define EHIP(expr)                                                                     \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return efail(ens, PG_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

int key_bits(int64_t n) {
    int b = 1;
    while (b < 31 && (int64_t(1) << b) < n) ++b;
    return b;
}

Batch batch_of(const pg_ensemble* ens, int y) {
    Batch b;
    b.sims = ens->dev_tab + (int64_t)y * ens->cfg.n_sims;
    b.n_sims = ens->cfg.n_sims;
    b.year = ens->cfg.first_year + y;
    b.n_tiles = ens->year_tiles[y];
    b.n_new = ens->year_new[y];
    return b;
}

// Enqueues growth year y for every simulation (no host synchronisation, no
// memset: capturable).
void enqueue_year(pg_ensemble* ens, int y, hipStream_t s, hipEvent_t* bfs_ev = nullptr) {
    pg_ctx* c = ens->ctx;
    const Batch b = batch_of(ens, y);
    const int m = ens->cfg.m;
    const int64_t TOT = ens->TOT;
    const int64_t Dn = b.n_new;
    unsigned long long* cnt = ens->counters;
    if (Dn <= 0) return;
    launch_birth_batch(s, b, c->birth);
    launch_weights_cdf_batch(s, b, c->in_deg, c->birth, c->logfit, c->w, c->cdf, c->scan_agg);
    const int lanes = ens->pa_lanes > 0 ? ens->pa_lanes : pa_draw_lanes(m);
    launch_pa_draw_batch(s, b, m, (uint32_t)y, lanes, c->cdf, c->cit, c->kind, c->cnt, c->near,
                         c->anext);
    bool any_nb = false;
    for (int si = 0; si < ens->cfg.n_sims; ++si) any_nb |= ens->sims[si].n_nb > 0;
    if (any_nb) {
        launch_root_dedup(s, TOT, Dn, m, c->cit, c->cnt, c->root_bits, c->r_wcount, c->r_wprefix,
                          c->scratch, c->r_excl + Dn, c->root_slot, c->uroot);
        int32_t* big_count = reinterpret_cast<int32_t*>(cnt + 3);
        launch_fill_i32(s, big_count, 0, 2);
        const CsrView g{c->fwd_off, c->fwd_adj, c->bwd_off[c->cur], c->bwd_adj[c->cur], TOT};
        if (bfs_ev) (void)hipEventRecord(bfs_ev[0], s);
        launch_bfs_root(s, g, c->w, c->uroot, c->r_excl + Dn, Dn, c->cfg.bfs_cap, c->S_buf,
                        c->Q_buf, c->Q_idx, c->ns_buf, c->big_list, big_count, cnt + 2,
                        bfs_stage_view(c->stage, c->tune_bfs_mode, c->tune_small_max,
                                       c->tune_mid_max),
                        c->byte_slots);
        if (bfs_ev) (void)hipEventRecord(bfs_ev[1], s);
        launch_nb_draw_batch(s, b, m, c->cfg.bfs_cap, (uint32_t)y, c->cdf, c->root_slot,
                             c->S_buf, c->Q_buf, c->Q_idx, c->ns_buf, c->cit, c->kind, c->cnt, c->near,
                             c->anext);
    }
    launch_sum_i32(s, c->near, Dn, cnt + 0);
    launch_sum_i32(s, c->anext, Dn, cnt + 1);
    launch_excl_scan_i32(s, c->cnt, Dn, c->new_off, c->scratch);
    launch_acc_i64(s, c->new_off + Dn, cnt + 4);
    launch_pa_compact_batch(s, b, m, c->cit, c->cnt, c->new_off, c->fwd_off, c->fwd_adj,
                            c->src_of, ens->stage);
    const int64_t UB = Dn * m;
    launch_fill_tail_u32(s, reinterpret_cast<uint32_t*>(ens->stage), c->new_off + Dn, UB,
                         (uint32_t)TOT);
    launch_radix_sort_pairs(s, reinterpret_cast<const uint32_t*>(ens->stage),
                            reinterpret_cast<const uint32_t*>(c->src_of), c->k_sorted, c->v_sorted,
                            c->k_tmp, c->v_tmp, UB, key_bits(TOT + 1), c->scratch);
    const int nxt = 1 - c->cur;
    launch_bwd_offsets_sorted(s, TOT, 0, c->bwd_off[c->cur], c->k_sorted, UB, c->bwd_off[nxt],
                              c->in_deg);
    launch_bwd_merge(s, TOT, TOT, c->bwd_off[c->cur], c->bwd_adj[c->cur], c->bwd_off[nxt],
                     c->v_sorted, c->bwd_adj[nxt], -1, ens->ETOT, c->scratch, c->cfg.max_papers + 2,
                     c->tune_merge_out);
    c->cur = nxt;
}

void drop_graph(pg_ensemble* ens) {
    if (ens->exec) (void)hipGraphExecDestroy(ens->exec);
    if (ens->graph) (void)hipGraphDestroy(ens->graph);
    ens->exec = nullptr;
    ens->graph = nullptr;
}

}  // namespace

extern "C" {

const char* pg_ensemble_error(const pg_ensemble* ens) {
    return ens ? ens->err.c_str() : g_ens_err.c_str();
}

int pg_ensemble_create(const pg_ensemble_config* cfg, const pg_sim_params* sims,
                       pg_ensemble** out) {
    g_ens_err.clear();
    pg_ensemble* ens = nullptr;
    if (!cfg || !sims || !out || cfg->size != sizeof(pg_ensemble_config))
        return efail(ens, PG_EINVAL, "pg_ensemble_create: null argument or struct size mismatch");
    *out = nullptr;
    if (cfg->n_sims < 1 || cfg->n_years < 0 || cfg->m < 1 || cfg->m > kMaxCites ||
        cfg->bfs_cap < 1 || cfg->bfs_cap > kBfsCapMax || !(cfg->growth_rate >= 0.0))
        return efail(ens, PG_EINVAL, "pg_ensemble_create: bad configuration");
    for (int s = 0; s < cfg->n_sims; ++s) {
        const pg_sim_params& p = sims[s];
        if (p.size != sizeof(pg_sim_params) || p.n0 < 1 || p.seed_edges < 0 || p.n_nb < 0 ||
            p.n_nb > cfg->m - 1 || !(p.inv_tau >= 0.0f) || !(p.alpha == p.alpha))
            return efail(ens, PG_EINVAL, "pg_ensemble_create: bad simulation parameters (sim " +
                                             std::to_string(s) + ")");
    }
    ens = new (std::nothrow) pg_ensemble();
    if (!ens) return efail(nullptr, PG_ENOMEM, "pg_ensemble_create: out of host memory");
    ens->cfg = *cfg;
    ens->sims.assign(sims, sims + cfg->n_sims);
    const int S = cfg->n_sims, Y = cfg->n_years, m = cfg->m;
    ens->R.resize(S); ens->RS.resize(S); ens->EB.resize(S); ens->EC.resize(S);
    ens->nfinal.resize(S); ens->D.assign(S, std::vector<int64_t>(Y, 0));
    int64_t R = 0, EB = 0;
    for (int s = 0; s < S; ++s) {
        int64_t n = sims[s].n0, tot_new = 0;
        for (int y = 0; y < Y; ++y) {
            const int64_t d = pg_growth_count(n, cfg->growth_rate);
            ens->D[s][y] = d;
            n += d;
            tot_new += d;
        }
        ens->nfinal[s] = n;
        ens->R[s] = R;
        ens->RS[s] = (n + 1 + kTile - 1) / kTile * kTile;
        R += ens->RS[s];
        ens->EB[s] = EB;
        ens->EC[s] = sims[s].seed_edges + tot_new * m;
        EB += ens->EC[s];
    }
    ens->TOT = R;
    ens->ETOT = EB;
    ens->year_new.assign(Y, 0);
    ens->year_tiles.assign(Y, 0);
    ens->host_tab.resize((size_t)Y * S);
    int64_t max_new = 1;
    for (int y = 0; y < Y; ++y) {
        int64_t nf = 0, tf = 0;
        for (int s = 0; s < S; ++s) {
            int64_t n = sims[s].n0;
            for (int k = 0; k < y; ++k) n += ens->D[s][k];
            BatchSim& b = ens->host_tab[(size_t)y * S + s];
            b.base = ens->R[s];
            b.n_exist = n;
            b.new_first = nf;
            b.n_new = ens->D[s][y];
            b.tile_first = tf;
            b.seed = sims[s].seed;
            b.alpha = sims[s].alpha;
            b.inv_tau = sims[s].inv_tau;
            b.n_pa = m - sims[s].n_nb;
            b.n_nb = sims[s].n_nb;
            nf += b.n_new;
            tf += (n + kTile - 1) / kTile;
        }
        ens->year_new[y] = nf;
        ens->year_tiles[y] = tf;
        max_new = std::max(max_new, nf);
    }
    pg_config c{};
    c.size = sizeof c;
    c.device = cfg->device;
    c.max_papers = ens->TOT;
    c.max_edges = ens->ETOT;
    c.max_new_per_year = max_new;
    c.max_cites = m;
    c.bfs_cap = cfg->bfs_cap;
    c.seed = 0;
    int rc = pg_ctx_create(&c, &ens->ctx);
    if (rc) {
        std::string m2 = pg_last_error(nullptr);
        delete ens;
        return efail(nullptr, rc, "pg_ensemble_create: " + m2);
    }
    pg_ctx* ctx = ens->ctx;
    hipError_t r = dalloc(&ens->dev_tab, std::max<int64_t>((int64_t)Y * S, 1));
    if (r == hipSuccess) r = dalloc(&ens->stage, max_new * m);
    if (r == hipSuccess) r = dalloc(&ens->counters, 8);
    if (r == hipSuccess && Y > 0)
        r = hipMemcpy(ens->dev_tab, ens->host_tab.data(), sizeof(BatchSim) * (size_t)Y * S,
                      hipMemcpyHostToDevice);
    if (r == hipSuccess) {
        int rc2 = ensure_nb_buffers(ctx);
        if (rc2) r = hipErrorOutOfMemory;
    }
    if (r == hipSuccess) r = hipMalloc(reinterpret_cast<void**>(&ctx->logfit), ens->TOT * 4);
    if (r != hipSuccess) {
        pg_ensemble_destroy(ens);
        return efail(nullptr, PG_ENOMEM, std::string("pg_ensemble_create: device allocation: ") +
                                             hipGetErrorString(r));
    }
    ctx->n = ens->TOT;
    ctx->e = ens->ETOT;
    ctx->n_logfit = ens->TOT;
    ctx->cur = 0;
    *out = ens;
    return PG_OK;
}

void pg_ensemble_destroy(pg_ensemble* ens) {
    if (!ens) return;
    drop_graph(ens);
    if (ens->dev_tab) (void)hipFree(ens->dev_tab);
    if (ens->stage) (void)hipFree(ens->stage);
    if (ens->counters) (void)hipFree(ens->counters);
    if (ens->ctx) pg_ctx_destroy(ens->ctx);
    delete ens;
}

int pg_ensemble_load(pg_ensemble* ens, int32_t sim, int64_t n, const int32_t* birth,
                     const float* logfit, int64_t n_edges, const int32_t* src,
                     const int32_t* dst) {
    if (!ens) return PG_EINVAL;
    if (sim != ens->loaded_upto)
        return efail(ens, PG_ESTATE, "pg_ensemble_load: simulations load in order 0, 1, ...");
    const pg_sim_params& p = ens->sims[sim];
    if (n != p.n0 || !birth || !logfit || n_edges < 0 || (n_edges > 0 && (!src || !dst)))
        return efail(ens, PG_EINVAL, "pg_ensemble_load: n must equal the simulation's n0");
    for (int64_t k = 0; k < n_edges; ++k) {
        if (src[k] < 0 || src[k] >= n || dst[k] < 0 || dst[k] >= n || src[k] == dst[k])
            return efail(ens, PG_EINVAL, "pg_ensemble_load: edge endpoint out of range");
    }
    // forward rows sorted and deduplicated, backward rows = transpose
    std::vector<int64_t> off(n + 1, 0);
    for (int64_t k = 0; k < n_edges; ++k) off[src[k] + 1]++;
    for (int64_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<int32_t> adj(n_edges);
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        for (int64_t k = 0; k < n_edges; ++k) adj[cur[src[k]]++] = dst[k];
    }
    const int64_t Rs = ens->R[sim], EBs = ens->EB[sim], RSs = ens->RS[sim];
    std::vector<int32_t> fo(RSs + 1);
    int64_t ne = 0;
    for (int64_t v = 0; v < n; ++v) {
        int32_t* b = adj.data() + off[v];
        int32_t* e = adj.data() + off[v + 1];
        std::sort(b, e);
        int32_t* u = std::unique(b, e);
        fo[v] = (int32_t)(EBs + ne);
        for (int32_t* q = b; q < u; ++q) adj[ne++] = *q + (int32_t)Rs;
    }
    if (ne > p.seed_edges)
        return efail(ens, PG_ECAP, "pg_ensemble_load: more distinct edges than seed_edges");
    for (int64_t v = n; v <= RSs; ++v) fo[v] = (int32_t)(EBs + ne);
    std::vector<int32_t> indeg(RSs, 0), bo(RSs + 1, 0), badj(ne);
    for (int64_t k = 0; k < ne; ++k) indeg[adj[k] - Rs]++;
    bo[0] = (int32_t)ens->bwd_fill;
    for (int64_t v = 0; v < RSs; ++v) bo[v + 1] = bo[v] + indeg[v];
    {
        std::vector<int32_t> cur(bo.begin(), bo.end() - 1);
        for (int64_t v = 0; v < n; ++v)
            for (int32_t k = fo[v] - (int32_t)EBs; k < fo[v + 1] - (int32_t)EBs; ++k)
                badj[cur[adj[k] - Rs]++ - ens->bwd_fill] = (int32_t)(v + Rs);
    }
    std::vector<int32_t> bfull(RSs, 0);
    std::copy(birth, birth + n, bfull.begin());
    pg_ctx* c = ens->ctx;
    EHIP(hipSetDevice(ens->cfg.device));
    EHIP(hipMemcpy(c->birth + Rs, bfull.data(), RSs * 4, hipMemcpyHostToDevice));
    EHIP(hipMemcpy(c->logfit + Rs, logfit, ens->nfinal[sim] * 4, hipMemcpyHostToDevice));
    EHIP(hipMemcpy(c->in_deg + Rs, indeg.data(), RSs * 4, hipMemcpyHostToDevice));
    EHIP(hipMemcpy(c->fwd_off + Rs, fo.data(), (RSs + 1) * 4, hipMemcpyHostToDevice));
    if (ne) EHIP(hipMemcpy(c->fwd_adj + EBs, adj.data(), ne * 4, hipMemcpyHostToDevice));
    EHIP(hipMemcpy(c->bwd_off[0] + Rs, bo.data(), (RSs + 1) * 4, hipMemcpyHostToDevice));
    if (ne) EHIP(hipMemcpy(c->bwd_adj[0] + ens->bwd_fill, badj.data(), ne * 4,
                           hipMemcpyHostToDevice));
    ens->bwd_fill += ne;
    ens->loaded_upto++;
    if (ens->loaded_upto == ens->cfg.n_sims) {
        c->loaded = true;
        c->cur = 0;
        int rc = pg_snapshot_save(c);
        if (rc) return efail(ens, rc, std::string("pg_ensemble_load: ") + c->err);
        ens->ready = true;
        ens->fresh = true;
    }
    return PG_OK;
}

int pg_ensemble_generate(pg_ensemble* ens, int32_t year_lo, int32_t year_hi, double mean_cites,
                         double mu, double sigma) {
    if (!ens) return PG_EINVAL;
    if (ens->loaded_upto != 0)
        return efail(ens, PG_ESTATE, "pg_ensemble_generate: simulations already loaded");
    const int64_t span = (int64_t)year_hi - year_lo + 1;
    if (span < 1 || span > 4096 || !(mean_cites >= 0.0) || !(sigma >= 0.0) || !std::isfinite(mu))
        return efail(ens, PG_EINVAL, "pg_ensemble_generate: bad year range or law parameters");
    const int S = ens->cfg.n_sims;
    std::vector<GenSim> tab(S);
    int64_t n_flat = 0, n_lf = 0;
    for (int s = 0; s < S; ++s) {
        tab[s] = GenSim{ens->R[s], n_flat, ens->sims[s].n0, n_lf, ens->EB[s], ens->sims[s].seed};
        n_flat += ens->sims[s].n0;
        n_lf += ens->nfinal[s];
    }
    pg_ctx* c = ens->ctx;
    EHIP(hipSetDevice(ens->cfg.device));
    hipStream_t st = c->stream;
    GenSim* d_tab = nullptr;
    int32_t *hist = nullptr, *cnt = nullptr, *zero = nullptr;
    int64_t *per_year = nullptr, *off = nullptr, *scan_scr = nullptr, *sim_edges = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr, *k_tmp = nullptr, *v_tmp = nullptr, *k_sorted = nullptr;
    int64_t* rscr = nullptr;
    auto release_tmp = [&]() {
        void* ps[] = {d_tab, hist, cnt, zero, per_year, off, scan_scr, sim_edges,
                      keys, vals, k_tmp, v_tmp, k_sorted, rscr};
        for (void* q : ps)
            if (q) (void)hipFree(q);
    };
    auto bail = [&](int code, const std::string& msg) {
        (void)hipStreamSynchronize(st);
        release_tmp();
        return efail(ens, code, "pg_ensemble_generate: " + msg);
    };
    hipError_t r = dalloc(&d_tab, S);
    if (r == hipSuccess) r = dalloc(&hist, (int64_t)S * span);
    if (r == hipSuccess) r = dalloc(&per_year, (int64_t)S * (span + 1));
    if (r == hipSuccess) r = dalloc(&cnt, n_flat);
    if (r == hipSuccess) r = dalloc(&off, n_flat + 1);
    if (r == hipSuccess) r = dalloc(&scan_scr, scan_scratch_elems(n_flat));
    if (r == hipSuccess) r = dalloc(&sim_edges, 2 * S);
    if (r == hipSuccess) r = dalloc(&zero, 1);
    if (r != hipSuccess) return bail(PG_ENOMEM, "device allocation failed");
    r = hipMemcpy(d_tab, tab.data(), sizeof(GenSim) * S, hipMemcpyHostToDevice);
    if (r != hipSuccess) return bail(PG_EHIP, hipGetErrorString(r));
    // Cleared on st itself: it is non-blocking, so a null-stream memset
    // would not be ordered before the histogram kernel.
    launch_fill_i32(st, hist, 0, (int64_t)S * span);
    launch_fill_i32(st, zero, 0, 1);
    // birth-year histogram -> per-year prefix (host: S x span values)
    launch_gen_hist(st, d_tab, S, n_flat, (int32_t)span, hist);
    std::vector<int32_t> h((size_t)S * span);
    r = hipMemcpyAsync(h.data(), hist, h.size() * 4, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return bail(PG_EHIP, hipGetErrorString(r));
    std::vector<int64_t> py((size_t)S * (span + 1));
    for (int s = 0; s < S; ++s) {
        py[(size_t)s * (span + 1)] = 0;
        for (int64_t y = 0; y < span; ++y)
            py[(size_t)s * (span + 1) + y + 1] = py[(size_t)s * (span + 1) + y] + h[(size_t)s * span + y];
    }
    r = hipMemcpy(per_year, py.data(), py.size() * 8, hipMemcpyHostToDevice);
    if (r != hipSuccess) return bail(PG_EHIP, hipGetErrorString(r));
    // counts, offsets, per-simulation edge totals (checked against seed_edges)
    launch_gen_counts(st, d_tab, S, n_flat, (int32_t)span, year_lo, per_year, mean_cites,
                      std::exp(-mean_cites), c->birth, cnt);
    launch_excl_scan_i32(st, cnt, n_flat, off, scan_scr);
    launch_gen_sim_edges(st, d_tab, S, off, sim_edges);
    std::vector<int64_t> se(2 * (size_t)S);
    r = hipGetLastError();
    if (r == hipSuccess) r = hipMemcpyAsync(se.data(), sim_edges, se.size() * 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return bail(PG_EHIP, hipGetErrorString(r));
    int64_t E = 0;
    for (int s = 0; s < S; ++s) {
        if (se[2 * s + 1] > ens->sims[s].seed_edges)
            return bail(PG_ECAP, "simulation " + std::to_string(s) + " draws " +
                                     std::to_string(se[2 * s + 1]) + " seed citations, more than "
                                     "its seed_edges");
        E += se[2 * s + 1];
    }
    r = dalloc(&keys, E);
    if (r == hipSuccess) r = dalloc(&vals, E);
    if (r == hipSuccess) r = dalloc(&k_tmp, E);
    if (r == hipSuccess) r = dalloc(&v_tmp, E);
    if (r == hipSuccess) r = dalloc(&k_sorted, E);
    if (r == hipSuccess) r = dalloc(&rscr, radix_scratch_elems(E));
    if (r != hipSuccess) return bail(PG_ENOMEM, "device allocation for the transpose failed");
    // forward rows, padding, backward CSR (stable sort of (target, citer)),
    // in-degrees, log-fitness
    launch_gen_rows(st, d_tab, S, n_flat, (int32_t)span, year_lo, per_year, c->birth, cnt, off,
                    sim_edges, c->fwd_off, c->fwd_adj, keys, vals);
    launch_gen_pad(st, d_tab, S, ens->TOT, sim_edges, c->birth, c->fwd_off);
    if (E > 0)
        launch_radix_sort_pairs(st, keys, vals, k_sorted, reinterpret_cast<uint32_t*>(c->bwd_adj[0]),
                                k_tmp, v_tmp, E, key_bits(ens->TOT), rscr);
    launch_bwd_offsets_sorted(st, 0, ens->TOT, zero, k_sorted, E, c->bwd_off[0], c->in_deg);
    launch_gen_indeg(st, ens->TOT, c->bwd_off[0], c->in_deg);
    launch_gen_logfit(st, d_tab, S, n_lf, mu, sigma, c->logfit);
    r = hipGetLastError();
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return bail(PG_EHIP, hipGetErrorString(r));
    release_tmp();
    ens->bwd_fill = E;
    ens->loaded_upto = S;
    c->loaded = true;
    c->cur = 0;
    int rc = pg_snapshot_save(c);
    if (rc) return efail(ens, rc, std::string("pg_ensemble_generate: ") + c->err);
    ens->ready = true;
    ens->fresh = true;
    return PG_OK;
}

int pg_ensemble_restore(pg_ensemble* ens) {
    if (!ens) return PG_EINVAL;
    if (!ens->ready) return efail(ens, PG_ESTATE, "pg_ensemble_restore: not all simulations loaded");
    int rc = pg_snapshot_restore(ens->ctx);
    if (rc) return efail(ens, rc, std::string("pg_ensemble_restore: ") + ens->ctx->err);
    ens->ctx->n = ens->TOT;
    ens->ctx->e = ens->ETOT;
    ens->fresh = true;
    ens->next_year = 0;
    return PG_OK;
}

int pg_ensemble_step(pg_ensemble* ens, pg_ensemble_stats* stats) {
    if (!ens) return PG_EINVAL;
    if (stats && stats->size != sizeof(pg_ensemble_stats))
        return efail(ens, PG_EINVAL, "pg_ensemble_step: pg_ensemble_stats.size mismatch");
    if (!ens->ready) return efail(ens, PG_ESTATE, "pg_ensemble_step: not all simulations loaded");
    if (ens->next_year >= ens->cfg.n_years)
        return efail(ens, PG_ESTATE, "pg_ensemble_step: every growth year is done");
    pg_ctx* c = ens->ctx;
    EHIP(hipSetDevice(ens->cfg.device));
    hipStream_t s = c->stream;
    const int y = ens->next_year;
    EHIP(hipEventRecord(c->ev0, s));
    launch_fill_i32(s, reinterpret_cast<int32_t*>(ens->counters), 0, 16);
    bool any_nb = false;
    for (int si = 0; si < ens->cfg.n_sims; ++si) any_nb |= ens->sims[si].n_nb > 0;
    hipEvent_t* bev = any_nb && ens->year_new[y] > 0 ? c->pev[PG_K_BFS] : nullptr;
    enqueue_year(ens, y, s, bev);
    EHIP(hipGetLastError());
    EHIP(hipEventRecord(c->ev1, s));
    EHIP(hipEventSynchronize(c->ev1));
    ens->fresh = false;
    ens->next_year = y + 1;
    unsigned long long h[8];
    EHIP(hipMemcpy(h, ens->counters, sizeof h, hipMemcpyDeviceToHost));
    if (stats) {
        float ms = 0.0f;
        EHIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->edges_added = (int64_t)h[4];
        stats->near_boundary = (int64_t)h[0];
        stats->pa_attempts = (int64_t)h[1];
        stats->shortfall = ens->year_new[y] * ens->cfg.m - (int64_t)h[4];
        stats->device_ms = ms;
        float bms = 0.0f;
        if (bev) EHIP(hipEventElapsedTime(&bms, bev[0], bev[1]));
        stats->bfs_ms = bms;
        stats->bfs_bytes = (double)h[2];
    }
    return PG_OK;
}

int pg_ensemble_year_draws(pg_ensemble* ens, int32_t sim, int32_t* cit, int8_t* kind,
                           int32_t* cnt, int32_t* near_boundary, float* w, double* cdf) {
    if (!ens) return PG_EINVAL;
    if (sim < 0 || sim >= ens->cfg.n_sims || !ens->ready)
        return efail(ens, PG_EINVAL, "pg_ensemble_year_draws: bad simulation or not loaded");
    if (ens->next_year < 1)
        return efail(ens, PG_ESTATE, "pg_ensemble_year_draws: no growth year done");
    const int y = ens->next_year - 1, m = ens->cfg.m;
    const BatchSim& b = ens->host_tab[(size_t)y * ens->cfg.n_sims + sim];
    pg_ctx* c = ens->ctx;
    EHIP(hipSetDevice(ens->cfg.device));
    const int64_t D = b.n_new, N = b.n_exist;
    if (cit && D) {
        EHIP(hipMemcpy(cit, c->cit + b.new_first * m, D * m * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < D * m; ++k)
            if (cit[k] >= 0) cit[k] -= (int32_t)b.base;
    }
    if (kind && D) EHIP(hipMemcpy(kind, c->kind + b.new_first * m, D * m, hipMemcpyDeviceToHost));
    if (cnt && D) EHIP(hipMemcpy(cnt, c->cnt + b.new_first, D * 4, hipMemcpyDeviceToHost));
    if (near_boundary && D)
        EHIP(hipMemcpy(near_boundary, c->near + b.new_first, D * 4, hipMemcpyDeviceToHost));
    if (w && N) EHIP(hipMemcpy(w, c->w + b.base, N * 4, hipMemcpyDeviceToHost));
    if (cdf && N) EHIP(hipMemcpy(cdf, c->cdf + b.base, N * 8, hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_ensemble_run(pg_ensemble* ens, int32_t use_graph, pg_ensemble_stats* stats) {
    if (!ens) return PG_EINVAL;
    if (stats && stats->size != sizeof(pg_ensemble_stats))
        return efail(ens, PG_EINVAL, "pg_ensemble_run: pg_ensemble_stats.size mismatch");
    if (!ens->ready) return efail(ens, PG_ESTATE, "pg_ensemble_run: not all simulations loaded");
    if (!ens->fresh)
        return efail(ens, PG_ESTATE, "pg_ensemble_run: restore the seed state first");
    pg_ctx* c = ens->ctx;
    EHIP(hipSetDevice(ens->cfg.device));
    hipStream_t s = c->stream;
    EHIP(hipEventRecord(c->ev0, s));
    launch_fill_i32(s, reinterpret_cast<int32_t*>(ens->counters), 0, 16);
    if (use_graph) {
        if (!ens->exec) {
            EHIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int y = 0; y < ens->cfg.n_years; ++y) enqueue_year(ens, y, s);
            hipGraph_t g = nullptr;
            hipError_t r = hipStreamEndCapture(s, &g);
            if (r != hipSuccess)
                return efail(ens, PG_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
            ens->graph = g;
            r = hipGraphInstantiate(&ens->exec, g, nullptr, nullptr, 0);
            if (r != hipSuccess) {
                drop_graph(ens);
                return efail(ens, PG_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r));
            }
        }
        EHIP(hipGraphLaunch(ens->exec, s));
        c->cur = ens->cfg.n_years % 2;
    } else {
        for (int y = 0; y < ens->cfg.n_years; ++y) enqueue_year(ens, y, s);
        EHIP(hipGetLastError());
    }
    EHIP(hipEventRecord(c->ev1, s));
    EHIP(hipEventSynchronize(c->ev1));
    ens->fresh = false;
    ens->next_year = ens->cfg.n_years;
    unsigned long long h[8];
    EHIP(hipMemcpy(h, ens->counters, sizeof h, hipMemcpyDeviceToHost));
    if (stats) {
        float ms = 0.0f;
        EHIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        int64_t slots = 0;
        for (int y = 0; y < ens->cfg.n_years; ++y) slots += ens->year_new[y] * ens->cfg.m;
        stats->edges_added = (int64_t)h[4];
        stats->near_boundary = (int64_t)h[0];
        stats->pa_attempts = (int64_t)h[1];
        stats->shortfall = slots - (int64_t)h[4];
        stats->device_ms = ms;
        stats->bfs_ms = 0.0;
        stats->bfs_bytes = (double)h[2];
    }
    return PG_OK;
}

int pg_ensemble_sim_size(pg_ensemble* ens, int32_t sim, int64_t* n_papers, int64_t* n_edges) {
    if (!ens) return PG_EINVAL;
    if (sim < 0 || sim >= ens->cfg.n_sims || !ens->ready)
        return efail(ens, PG_EINVAL, "pg_ensemble_sim_size: bad simulation or not loaded");
    int64_t N = ens->sims[sim].n0;
    for (int y = 0; y < ens->next_year; ++y) N += ens->D[sim][y];
    int32_t b0 = 0, b1 = 0;
    EHIP(hipMemcpy(&b0, ens->ctx->fwd_off + ens->R[sim], 4, hipMemcpyDeviceToHost));
    EHIP(hipMemcpy(&b1, ens->ctx->fwd_off + ens->R[sim] + N, 4, hipMemcpyDeviceToHost));
    if (n_papers) *n_papers = N;
    if (n_edges) *n_edges = b1 - b0;
    return PG_OK;
}

int pg_ensemble_download(pg_ensemble* ens, int32_t sim, int32_t* fwd_off, int32_t* fwd_adj,
                         int32_t* bwd_off, int32_t* bwd_adj, int32_t* in_deg, int32_t* birth) {
    int64_t N = 0, E = 0;
    int rc = pg_ensemble_sim_size(ens, sim, &N, &E);
    if (rc) return rc;
    pg_ctx* c = ens->ctx;
    const int64_t Rs = ens->R[sim];
    std::vector<int32_t> fo(N + 1), bo(N + 1);
    EHIP(hipMemcpy(fo.data(), c->fwd_off + Rs, (N + 1) * 4, hipMemcpyDeviceToHost));
    EHIP(hipMemcpy(bo.data(), c->bwd_off[c->cur] + Rs, (N + 1) * 4, hipMemcpyDeviceToHost));
    if (fwd_off)
        for (int64_t v = 0; v <= N; ++v) fwd_off[v] = fo[v] - fo[0];
    if (bwd_off)
        for (int64_t v = 0; v <= N; ++v) bwd_off[v] = bo[v] - bo[0];
    if (fwd_adj && E) {
        EHIP(hipMemcpy(fwd_adj, c->fwd_adj + fo[0], E * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < E; ++k) fwd_adj[k] -= (int32_t)Rs;
    }
    const int64_t EB = bo[N] - bo[0];
    if (bwd_adj && EB) {
        EHIP(hipMemcpy(bwd_adj, c->bwd_adj[c->cur] + bo[0], EB * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < EB; ++k) bwd_adj[k] -= (int32_t)Rs;
    }
    if (in_deg) EHIP(hipMemcpy(in_deg, c->in_deg + Rs, N * 4, hipMemcpyDeviceToHost));
    if (birth) EHIP(hipMemcpy(birth, c->birth + Rs, N * 4, hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_ensemble_set_tuning(pg_ensemble* ens, const pg_tuning* t) {
    if (!ens || !t || t->size != sizeof(pg_tuning))
        return efail(ens, PG_EINVAL, "pg_ensemble_set_tuning: null or struct size mismatch");
    if (const char* e = tuning_error(t))
        return efail(ens, PG_EINVAL, std::string("pg_ensemble_set_tuning: ") + e);
    pg_ctx* c = ens->ctx;
    if (t->pa_lanes != ens->pa_lanes || t->bfs_mode != c->tune_bfs_mode ||
        t->bfs_small_max != c->tune_small_max || t->bfs_mid_max != c->tune_mid_max ||
        t->merge_out != c->tune_merge_out)
        drop_graph(ens);  // the captured graph holds the old launches
    ens->pa_lanes = t->pa_lanes;
    c->tune_bfs_mode = t->bfs_mode;
    c->tune_small_max = t->bfs_small_max;
    c->tune_mid_max = t->bfs_mid_max;
    c->tune_merge_out = t->merge_out;
    return PG_OK;
}

}  // extern "C"

namespace {

double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n == 0 ? 0.0 : (n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]));
}

}  // namespace

extern "C" {

int pg_ensemble_autotune(pg_ensemble* ens, int32_t reps, pg_tune_report* report) {
    if (!ens) return PG_EINVAL;
    if (report && report->size != sizeof(pg_tune_report))
        return efail(ens, PG_EINVAL, "pg_ensemble_autotune: pg_tune_report.size mismatch");
    if (reps < 1 || reps > 64) return efail(ens, PG_EINVAL, "pg_ensemble_autotune: reps in [1, 64]");
    if (!ens->ready) return efail(ens, PG_ESTATE, "pg_ensemble_autotune: not all simulations loaded");
    std::vector<pg_tuning> cand;
    std::vector<double> ms;
    pg_ensemble_stats ref{};
    bool have_ref = false;
    // one candidate: untimed run (capture), then reps timed runs from the seed state
    auto trial = [&](const pg_tuning& t) -> int {
        int rc = pg_ensemble_set_tuning(ens, &t);
        if (rc) return rc;
        std::vector<double> v;
        for (int r = 0; r <= reps; ++r) {
            if ((rc = pg_ensemble_restore(ens))) return rc;
            pg_ensemble_stats st{};
            st.size = sizeof st;
            if ((rc = pg_ensemble_run(ens, 1, &st))) return rc;
            if (!have_ref) { ref = st; have_ref = true; }
            if (st.edges_added != ref.edges_added || st.near_boundary != ref.near_boundary ||
                st.pa_attempts != ref.pa_attempts)
                return efail(ens, PG_ESTATE, "pg_ensemble_autotune: a launch shape changed the run");
            if (r > 0) v.push_back(st.device_ms);
        }
        cand.push_back(t);
        ms.push_back(median_of(v));
        return PG_OK;
    };
    auto shape = [](int lanes, int mode) {
        pg_tuning t{};
        t.size = sizeof t;
        t.pa_lanes = lanes;
        t.bfs_mode = mode;
        return t;
    };
    const int lanes_c[] = {8, 16, 32, 64};
    const int def_lanes = pa_draw_lanes(ens->cfg.m);
    double default_ms = 0.0;
    for (int l : lanes_c) {
        if (int rc = trial(shape(l, 0))) return rc;
        if (l == def_lanes) default_ms = ms.back();
    }
    size_t best = std::min_element(ms.begin(), ms.end()) - ms.begin();
    const int best_lanes = cand[best].pa_lanes;
    for (int mode : {1, 2})
        if (int rc = trial(shape(best_lanes, mode))) return rc;
    best = std::min_element(ms.begin(), ms.end()) - ms.begin();
    // then the wavefront tiers' limits and the merge tile, one knob at a
    // time from the best so far
    auto vary = [&](int which, std::initializer_list<int> vals) -> int {
        for (int v : vals) {
            pg_tuning t = cand[best];
            (which == 0 ? t.bfs_small_max : which == 1 ? t.bfs_mid_max : t.merge_out) = v;
            if (int rc = trial(t)) return rc;
        }
        best = std::min_element(ms.begin(), ms.end()) - ms.begin();
        return PG_OK;
    };
    if (int rc = vary(0, {512, 1024})) return rc;
    if (int rc = vary(1, {2048})) return rc;
    if (int rc = vary(2, {512, 2048})) return rc;
    const pg_tuning chosen = cand[best];
    if (int rc = pg_ensemble_set_tuning(ens, &chosen)) return rc;
    if (int rc = pg_ensemble_restore(ens)) return rc;
    if (report) {
        report->n_trials = (int32_t)std::min<size_t>(cand.size(), PG_TUNE_MAX_TRIALS);
        for (int k = 0; k < report->n_trials; ++k) {
            report->trial[k] = cand[k];
            report->trial_ms[k] = ms[k];
        }
        report->chosen = chosen;
        report->chosen_ms = ms[best];
        report->default_ms = default_ms;
        report->max_rel_dev = 0.0;
    }
    return PG_OK;
}

}  // extern "C"
