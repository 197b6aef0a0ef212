// pg_api.hip -- host implementation of the C ABI declared in
// include/paragraph.h: context and device arena, seed-graph load, the
// per-year step (replacing buildOneNodeConnections_timed,
// src/kernel.cu:1125-1586, and the year-boundary update,
// src/kernel.cu:1766-1815), downloads, and the stand-alone BFS.
//
// HBM layout per context (int32 ids and offsets; N = papers, E = edges):
//   birth, in_deg : int32[N_cap]        logfit : float[N_logfit]
//   fwd_off       : int32[N_cap+1]      fwd_adj : int32[E_cap]   (appended)
//   bwd_off x2    : int32[N_cap+1]      bwd_adj x2 : int32[E_cap] (ping-pong)
//   w : float[N_cap], cdf : double[N_cap]       (this year's PA weights/CDF)
//   cit/kind : [max_new * max_cites], cnt/near/anext : [max_new]
//   add : int32[N_cap] (per-target new citations), src_of + radix buffers
//   sized for one year's edges, and an int64 scratch for scans.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_ctx.h"
#include "pg_internal.h"

using namespace pg;


namespace {

std::string g_create_err;

void release(pg_ctx* c) {
    void* ptrs[] = {c->birth, c->in_deg, c->logfit, c->fwd_off, c->fwd_adj, c->bwd_off[0],
                    c->bwd_off[1], c->bwd_adj[0], c->bwd_adj[1], c->w, c->cdf, c->cdf_idx,
                    c->scan_agg,
                    c->cit, c->kind, c->cnt,
                    c->near, c->anext, c->new_off, c->src_of, c->k_sorted,
                    c->v_sorted, c->k_tmp, c->v_tmp, c->scratch, c->d_sum, c->d_tmp_i32,
                    c->s_birth, c->s_in_deg, c->s_fwd_off, c->s_fwd_adj, c->s_bwd_off,
                    c->s_bwd_adj, c->S_buf, c->Q_buf, c->Q_idx, c->ns_buf, c->byte_slots, c->root_bits, c->r_wcount,
                    c->r_wprefix, c->r_excl,
                    c->root_slot, c->uroot, c->big_list, c->stage.buf, c->stage.busy};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_sum) (void)hipHostFree(c->h_sum);
    for (int k = 0; k < PG_K_COUNT; ++k)
        for (int j = 0; j < 2; ++j) {
            if (c->pev[k][j]) (void)hipEventDestroy(c->pev[k][j]);
            if (c->pev2[k][j]) (void)hipEventDestroy(c->pev2[k][j]);
        }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    seq_release(c);
    abm_release(c);
}

int64_t scratch_need(int64_t max_papers, int64_t max_new, int64_t year_edges) {
    int64_t a = (max_papers + 2) + scan_scratch_elems(max_papers + 1);
    int64_t b = scan_scratch_elems(max_new + 1);
    int64_t c = radix_scratch_elems(year_edges);
    return std::max(a, std::max(b, c));
}

}  // namespace

// Grows the per-year edge buffers to hold `need` edges.  The new buffers are
// allocated first and the old ones freed only after all of them succeeded,
// so a failed growth leaves the context with its old, consistent buffers.
int ensure_year_edges(pg_ctx* ctx, int64_t need) {
    if (need <= ctx->year_edge_cap) return PG_OK;
    const int64_t cap = std::max(need, ctx->year_edge_cap * 2);
    int32_t* src_of = nullptr;
    uint32_t *k_sorted = nullptr, *v_sorted = nullptr, *k_tmp = nullptr, *v_tmp = nullptr;
    int64_t* scratch = nullptr;
    hipError_t r = dalloc(&src_of, cap);
    if (r == hipSuccess) r = dalloc(&k_sorted, cap);
    if (r == hipSuccess) r = dalloc(&v_sorted, cap);
    if (r == hipSuccess) r = dalloc(&k_tmp, cap);
    if (r == hipSuccess) r = dalloc(&v_tmp, cap);
    if (r == hipSuccess)
        r = dalloc(&scratch, scratch_need(ctx->cfg.max_papers, ctx->cfg.max_new_per_year, cap));
    if (r != hipSuccess) {
        void* got[] = {src_of, k_sorted, v_sorted, k_tmp, v_tmp, scratch};
        for (void* q : got)
            if (q) (void)hipFree(q);
        return fail(ctx, r == hipErrorOutOfMemory ? PG_ENOMEM : PG_EHIP,
                    std::string("per-year edge buffers for ") + std::to_string(cap) +
                        " edges: " + hipGetErrorString(r));
    }
    // the old buffers may still be read by queued work
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    void* old[] = {ctx->src_of, ctx->k_sorted, ctx->v_sorted, ctx->k_tmp, ctx->v_tmp, ctx->scratch};
    for (void* q : old)
        if (q) (void)hipFree(q);
    ctx->src_of = src_of;
    ctx->k_sorted = k_sorted;
    ctx->v_sorted = v_sorted;
    ctx->k_tmp = k_tmp;
    ctx->v_tmp = v_tmp;
    ctx->scratch = scratch;
    ctx->year_edge_cap = cap;
    return PG_OK;
}

int key_bits_for(int64_t n) {
    int b = 1;
    while (b < 31 && (int64_t(1) << b) < n) ++b;
    return b;
}


CsrView view(const pg_ctx* c) {
    return CsrView{c->fwd_off, c->fwd_adj, c->bwd_off[c->cur], c->bwd_adj[c->cur], c->n};
}

int ensure_bfs_stage(pg_ctx* ctx) {
    if (ctx->stage.buf) return PG_OK;
    const int slots = bfs_stage_slots(ctx->cfg.device);
    if (slots <= 0) return PG_OK;  // no pool: every root takes the unstaged sweep
    BfsStage st;
    hipError_t r = dalloc(&st.buf, (int64_t)slots * kBfsStageEntries);
    if (r == hipSuccess) r = dalloc(&st.busy, slots);
    if (r != hipSuccess) {
        (void)hipFree(st.buf); (void)hipFree(st.busy);
        return fail(ctx, PG_ENOMEM, "2-hop staging pool: device allocation failed");
    }
    // Cleared on the context stream: a null-stream hipMemset is not ordered
    // with the non-blocking context stream, so the year's kernels could see
    // the flags before the clear lands.
    launch_fill_i32(ctx->stream, st.busy, 0, slots);
    st.n_slots = slots;
    st.cap_entries = kBfsStageEntries;
    ctx->stage = st;
    return PG_OK;
}

int ensure_nb_buffers(pg_ctx* ctx) {
    if (ctx->S_buf) return PG_OK;
    if (int rc = ensure_bfs_stage(ctx)) return rc;
    const int64_t D = std::max<int64_t>(ctx->cfg.max_new_per_year, 1);
    const int64_t cap = ctx->cfg.bfs_cap;
    hipError_t r = dalloc(&ctx->S_buf, D * cap);
    if (r == hipSuccess) r = dalloc(&ctx->Q_buf, D * cap);
    if (r == hipSuccess) r = dalloc(&ctx->Q_idx, D * kQIdx);
    if (r == hipSuccess) r = dalloc(&ctx->ns_buf, D);
    if (r == hipSuccess) r = dalloc(&ctx->byte_slots, (int64_t)kByteSlots * kByteSlotStride);
    const int64_t nw = ctx->cfg.max_papers / 32 + 1;
    if (r == hipSuccess) r = dalloc(&ctx->root_bits, nw);
    if (r == hipSuccess) r = dalloc(&ctx->r_wcount, nw);
    if (r == hipSuccess) r = dalloc(&ctx->r_wprefix, nw + 1);
    if (r == hipSuccess) r = dalloc(&ctx->r_excl, D + 1);
    if (r == hipSuccess) r = dalloc(&ctx->root_slot, D);
    if (r == hipSuccess) r = dalloc(&ctx->uroot, D);
    if (r == hipSuccess) r = dalloc(&ctx->big_list, 2 * D);
    if (r != hipSuccess)
        return fail(ctx, PG_ENOMEM, "neighbourhood buffers: device allocation failed");
    // Zeroed on the context stream, ahead of the kernels that use them (this
    // runs between launches of a year step, and that stream does not wait on
    // the null stream).
    launch_fill_i32(ctx->stream, reinterpret_cast<int32_t*>(ctx->byte_slots), 0,
                    2 * (int64_t)kByteSlots * kByteSlotStride);
    launch_fill_i32(ctx->stream, reinterpret_cast<int32_t*>(ctx->root_bits), 0, nw);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

extern "C" {

int pg_version(void) { return PG_ABI_VERSION; }

const char* pg_last_error(const pg_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int pg_ctx_create(const pg_config* cfg, pg_ctx** out) {
    g_create_err.clear();
    if (!cfg || !out || cfg->size != sizeof(pg_config)) {
        g_create_err = "pg_ctx_create: null argument or pg_config.size mismatch";
        return PG_EINVAL;
    }
    *out = nullptr;
    if (cfg->max_papers < 1 || cfg->max_papers >= (int64_t(1) << 31) - 1 || cfg->max_edges < 0 ||
        cfg->max_edges >= (int64_t(1) << 31) - 1 || cfg->max_new_per_year < 0 ||
        cfg->max_cites < 0 || cfg->max_cites > kMaxCites || cfg->bfs_cap < 1 ||
        cfg->bfs_cap > kBfsCapMax) {
        g_create_err = "pg_ctx_create: capacity out of range (papers/edges < 2^31-1, "
                       "max_cites <= 64, 1 <= bfs_cap <= 4096)";
        return PG_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        g_create_err = "pg_ctx_create: no HIP device with that ordinal";
        return PG_ENODEV;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_err = std::string("pg_ctx_create: device is not gfx950 (") + prop.gcnArchName + ")";
        return PG_ENODEV;
    }
    pg_ctx* ctx = new (std::nothrow) pg_ctx();
    if (!ctx) { g_create_err = "pg_ctx_create: out of host memory"; return PG_ENOMEM; }
    ctx->cfg = *cfg;
    auto bail = [&](int code) {
        g_create_err = ctx->err;
        release(ctx);
        delete ctx;
        return code;
    };
#define CREATE_HIP(expr)                                                       \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);      \
            return bail(_e == hipErrorOutOfMemory ? PG_ENOMEM : PG_EHIP);      \
        }                                                                      \
    } while (0)
    CREATE_HIP(hipSetDevice(cfg->device));
    CREATE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    CREATE_HIP(hipEventCreate(&ctx->ev0));
    CREATE_HIP(hipEventCreate(&ctx->ev1));
    const int64_t N = cfg->max_papers, E = std::max<int64_t>(cfg->max_edges, 1);
    const int64_t D = std::max<int64_t>(cfg->max_new_per_year, 1);
    const int64_t M = std::max<int64_t>(cfg->max_cites, 1);
    const int64_t tiles = (N + kTile - 1) / kTile + 1;
    CREATE_HIP(dalloc(&ctx->birth, N));
    CREATE_HIP(dalloc(&ctx->in_deg, N));
    CREATE_HIP(dalloc(&ctx->fwd_off, N + 1));
    CREATE_HIP(dalloc(&ctx->fwd_adj, E));
    for (int b = 0; b < 2; ++b) {
        CREATE_HIP(dalloc(&ctx->bwd_off[b], N + 1));
        CREATE_HIP(dalloc(&ctx->bwd_adj[b], E));
    }
    CREATE_HIP(dalloc(&ctx->w, N));
    CREATE_HIP(dalloc(&ctx->cdf, N));
    CREATE_HIP(dalloc(&ctx->cdf_idx, kCdfIdx + 2));
    CREATE_HIP(dalloc(&ctx->scan_agg, tiles));
    CREATE_HIP(dalloc(&ctx->cit, D * M));
    CREATE_HIP(dalloc(&ctx->kind, D * M));
    CREATE_HIP(dalloc(&ctx->cnt, D));
    CREATE_HIP(dalloc(&ctx->near, D));
    CREATE_HIP(dalloc(&ctx->anext, D));
    CREATE_HIP(dalloc(&ctx->new_off, D + 2));
    CREATE_HIP(dalloc(&ctx->d_sum, 5));
    CREATE_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->h_sum), 5 * sizeof(unsigned long long),
                             hipHostMallocDefault));
    for (int k = 0; k < PG_K_COUNT; ++k)
        for (int j = 0; j < 2; ++j) {
            CREATE_HIP(hipEventCreate(&ctx->pev[k][j]));
            CREATE_HIP(hipEventCreate(&ctx->pev2[k][j]));
        }
    {
        int rc = ensure_year_edges(ctx, D * M);
        if (rc) return bail(rc);
    }
    (void)noalloc_setup(ctx);
#undef CREATE_HIP
    *out = ctx;
    return PG_OK;
}

void pg_ctx_destroy(pg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->cfg.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
}

int pg_graph_load(pg_ctx* ctx, int64_t n, const int32_t* birth, const float* logfit,
                  int64_t n_logfit, int64_t n_edges, const int32_t* src, const int32_t* dst) {
    if (!ctx) return PG_EINVAL;
    if (n < 1 || !birth || !logfit || n_logfit < n || n_edges < 0 ||
        (n_edges > 0 && (!src || !dst)))
        return fail(ctx, PG_EINVAL, "pg_graph_load: bad arguments");
    if (n > ctx->cfg.max_papers) return fail(ctx, PG_ECAP, "pg_graph_load: n > max_papers");
    if (n_logfit > ctx->cfg.max_papers) n_logfit = ctx->cfg.max_papers;
    for (int64_t k = 0; k < n_edges; ++k) {
        if (src[k] < 0 || src[k] >= n || dst[k] < 0 || dst[k] >= n)
            return fail(ctx, PG_EINVAL, "pg_graph_load: edge endpoint out of range");
        if (src[k] == dst[k]) return fail(ctx, PG_EINVAL, "pg_graph_load: self-citation");
    }
    // forward CSR by counting sort on src, rows sorted and deduplicated
    std::vector<int64_t> off(n + 1, 0);
    for (int64_t k = 0; k < n_edges; ++k) off[src[k] + 1]++;
    for (int64_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<int32_t> adj(n_edges);
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        for (int64_t k = 0; k < n_edges; ++k) adj[cur[src[k]]++] = dst[k];
    }
    std::vector<int32_t> fo(n + 1, 0);
    int64_t out_e = 0;
    for (int64_t v = 0; v < n; ++v) {
        int32_t* b = adj.data() + off[v];
        int32_t* e = adj.data() + off[v + 1];
        std::sort(b, e);
        int32_t* u = std::unique(b, e);
        for (int32_t* p = b; p < u; ++p) adj[out_e++] = *p;
        fo[v + 1] = (int32_t)out_e;
    }
    adj.resize(out_e);
    if (out_e > ctx->cfg.max_edges) return fail(ctx, PG_ECAP, "pg_graph_load: edges > max_edges");
    // backward CSR: transpose, citers ascending
    std::vector<int32_t> bo(n + 1, 0), badj(out_e), indeg(n, 0);
    for (int64_t k = 0; k < out_e; ++k) indeg[adj[k]]++;
    for (int64_t v = 0; v < n; ++v) bo[v + 1] = bo[v] + indeg[v];
    {
        std::vector<int32_t> cur(bo.begin(), bo.end() - 1);
        for (int64_t v = 0; v < n; ++v)
            for (int32_t k = fo[v]; k < fo[v + 1]; ++k) badj[cur[adj[k]]++] = (int32_t)v;
    }
    PG_HIP(hipSetDevice(ctx->cfg.device));
    // the device state is overwritten below: until the load completes the
    // context holds no graph
    ctx->loaded = false;
    if (!ctx->logfit) PG_HIP(dalloc(&ctx->logfit, ctx->cfg.max_papers));
    hipStream_t s = ctx->stream;
    PG_HIP(hipMemcpyAsync(ctx->birth, birth, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->logfit, logfit, n_logfit * sizeof(float), hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->in_deg, indeg.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->fwd_off, fo.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (out_e)
        PG_HIP(hipMemcpyAsync(ctx->fwd_adj, adj.data(), out_e * sizeof(int32_t), hipMemcpyHostToDevice, s));
    ctx->cur = 0;
    PG_HIP(hipMemcpyAsync(ctx->bwd_off[0], bo.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (out_e)
        PG_HIP(hipMemcpyAsync(ctx->bwd_adj[0], badj.data(), out_e * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PG_HIP(hipStreamSynchronize(s));
    ctx->n = n;
    ctx->e = out_e;
    ctx->n_logfit = n_logfit;
    ctx->logfit_owned = true;
    ctx->loaded = true;
    ctx->last_model = -1;
    ctx->steps_done = 0;
    ctx->last_year = 0;
    ctx->last_year_index = -1;
    ctx->err.clear();
    return PG_OK;
}

int pg_year_step(pg_ctx* ctx, const pg_year_params* p, pg_year_stats* stats) {
    if (!ctx) return PG_EINVAL;
    if (!p || p->size != sizeof(pg_year_params) ||
        (stats && stats->size != sizeof(pg_year_stats)))
        return fail(ctx, PG_EINVAL, "pg_year_step: null params or struct size mismatch");
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_year_step: no graph loaded");
    const int64_t N = ctx->n, D = p->n_new;
    if (D < 0) return fail(ctx, PG_EINVAL, "pg_year_step: n_new < 0");
    if (D > ctx->cfg.max_new_per_year) return fail(ctx, PG_ECAP, "pg_year_step: n_new > max_new_per_year");
    if (N + D > ctx->cfg.max_papers) return fail(ctx, PG_ECAP, "pg_year_step: graph would exceed max_papers");
    if (N + D > ctx->n_logfit) return fail(ctx, PG_ECAP, "pg_year_step: log-fitness not supplied for the new papers");
    if (p->model == PG_MODEL_ABM) return abm_year_step(ctx, p, stats);
    if (p->model != PG_MODEL_PA && p->model != PG_MODEL_ER)
        return fail(ctx, PG_EINVAL, "pg_year_step: unknown model");
    if (p->model == PG_MODEL_PA) {
        if (p->m < 0 || p->m > ctx->cfg.max_cites)
            return fail(ctx, PG_ECAP, "pg_year_step: m exceeds max_cites");
        if (p->n_nb < 0 || (p->m > 0 && p->n_nb > p->m - 1) || (p->m == 0 && p->n_nb != 0))
            return fail(ctx, PG_EINVAL, "pg_year_step: n_nb must be in [0, m-1]");
        if (!(p->alpha == p->alpha) || !(p->inv_tau >= 0.0f))
            return fail(ctx, PG_EINVAL, "pg_year_step: alpha must be finite and inv_tau >= 0");
    } else if (!(p->er_p >= 0.0 && p->er_p <= 1.0)) {
        return fail(ctx, PG_EINVAL, "pg_year_step: er_p must be in [0,1]");
    }
    PG_HIP(hipSetDevice(ctx->cfg.device));
    {
        int rc = seq_record_snapshot(ctx, p->year, N, D, p->m, p->model);
        if (rc) return rc;
    }
    hipStream_t s = ctx->stream;
    const uint64_t seed = ctx->cfg.seed;
    const uint32_t yi = (uint32_t)p->year_index;
    bool used[PG_K_COUNT] = {};
    bool used2[PG_K_COUNT] = {};
    auto on = [&](int k) { return ctx->prof && ((ctx->prof_mask >> k) & 1u); };
    auto pbeg = [&](int k) { if (on(k)) (void)hipEventRecord(ctx->pev[k][0], s); };
    auto pend = [&](int k) { if (on(k)) { (void)hipEventRecord(ctx->pev[k][1], s); used[k] = true; } };
    // second interval of a group within the same step (NB_DRAW: dedup, then draws)
    auto pbeg2 = [&](int k) { if (on(k)) (void)hipEventRecord(ctx->pev2[k][0], s); };
    auto pend2 = [&](int k) { if (on(k)) { (void)hipEventRecord(ctx->pev2[k][1], s); used2[k] = true; } };
    // PA years whose bound D*m fits max_edges run without a host round trip
    // mid-year: the append sorts D*m keys (the unfilled tail padded with key
    // N+D, past every paper) and the merge reads its edge total on the device
    const int64_t UB = D * (int64_t)std::max(p->m, 0);
    const bool one_sync = p->model == PG_MODEL_PA && !ctx->seq_recording && D > 0 &&
                          ctx->e + UB <= ctx->cfg.max_edges;
    if (one_sync) {
        int rc = ensure_year_edges(ctx, UB);
        if (rc) return rc;
    }
    PG_HIP(hipEventRecord(ctx->ev0, s));
    int64_t e_year = 0, near_total = 0, attempts = 0;
    double bfs_bytes = 0.0;
    int64_t tot = 0;
    unsigned long long hs[4] = {0, 0, 0, 0};
    if (D > 0) {
        launch_fill_i32(s, ctx->birth + N, p->year, D);
        // [0] near, [1] attempts, [2] BFS bytes, [3] tier counts, [4] edges
        unsigned long long* sums = ctx->d_sum;
        launch_fill_i32(s, reinterpret_cast<int32_t*>(sums), 0, 10);
        if (p->model == PG_MODEL_PA) {
            const int m = p->m, n_nb = p->n_nb;
            pbeg(PG_K_WEIGHTS_CDF);
            launch_weights_cdf(s, N, ctx->in_deg, ctx->birth, ctx->logfit, p->year, p->alpha,
                               p->inv_tau, ctx->w, ctx->cdf, ctx->scan_agg, ctx->cdf_idx);
            pend(PG_K_WEIGHTS_CDF);
            pbeg(PG_K_PA_DRAW);
            launch_pa_draw(s, N, ctx->cdf, ctx->cdf_idx, D, m, m - n_nb, seed, yi, ctx->cit,
                           ctx->kind, ctx->cnt, ctx->near, ctx->anext, ctx->tune_pa_lanes);
            pend(PG_K_PA_DRAW);
            if (n_nb > 0) {
                {
                    int rc = ensure_nb_buffers(ctx);
                    if (rc) return rc;
                }
                pbeg(PG_K_NB_DRAW);
                launch_root_dedup(s, N, D, m, ctx->cit, ctx->cnt, ctx->root_bits, ctx->r_wcount,
                                  ctx->r_wprefix, ctx->scratch, ctx->r_excl + D, ctx->root_slot,
                                  ctx->uroot);
                pend(PG_K_NB_DRAW);
                int32_t* big_count = reinterpret_cast<int32_t*>(sums + 3);
                pbeg(PG_K_BFS);
                launch_bfs_root(s, view(ctx), ctx->w, ctx->uroot, ctx->r_excl + D, D,
                                ctx->cfg.bfs_cap, ctx->S_buf, ctx->Q_buf, ctx->Q_idx, ctx->ns_buf,
                                ctx->big_list, big_count, sums + 2,
                                bfs_stage_view(ctx->stage, ctx->tune_bfs_mode,
                                               ctx->tune_small_max, ctx->tune_mid_max),
                                ctx->byte_slots);
                pend(PG_K_BFS);
                pbeg2(PG_K_NB_DRAW);
                launch_nb_draw(s, N, D, m, n_nb, ctx->cfg.bfs_cap, seed, yi, ctx->cdf,
                               ctx->cdf_idx, ctx->root_slot, ctx->S_buf, ctx->Q_buf, ctx->Q_idx,
                               ctx->ns_buf, ctx->cit,
                               ctx->kind, ctx->cnt, ctx->near, ctx->anext);
                pend2(PG_K_NB_DRAW);
            }
            launch_sum_i32(s, ctx->near, D, sums + 0);
            launch_sum_i32(s, ctx->anext, D, sums + 1);
        } else {
            pbeg(PG_K_ER);
            launch_er_count(s, N, D, p->er_p, seed, yi, ctx->cnt);
        }
        launch_excl_scan_i32(s, ctx->cnt, D, ctx->new_off, ctx->scratch);
        PG_HIP(hipGetLastError());
        if (!one_sync) {
            launch_acc_i64(s, ctx->new_off + D, sums + 4);
            PG_HIP(hipMemcpyAsync(ctx->h_sum, sums, 5 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            PG_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < 4; ++k) hs[k] = ctx->h_sum[k];
            tot = (int64_t)ctx->h_sum[4];
            e_year = tot;
            if (ctx->e + e_year > ctx->cfg.max_edges)
                return fail(ctx, PG_ECAP, "pg_year_step: edges would exceed max_edges");
            int rc = ensure_year_edges(ctx, e_year);
            if (rc) return rc;
        }
        if (p->model == PG_MODEL_PA) {
            pbeg(PG_K_CSR_APPEND);
            launch_pa_compact(s, N, D, p->m, ctx->cit, ctx->cnt, ctx->new_off, ctx->e, ctx->fwd_off,
                              ctx->fwd_adj, ctx->src_of);
        } else {
            launch_er_write(s, N, D, p->er_p, seed, yi, ctx->new_off, ctx->e, ctx->fwd_off,
                            ctx->fwd_adj, ctx->src_of);
            pend(PG_K_ER);
            pbeg(PG_K_CSR_APPEND);
        }
        uint32_t* keys = reinterpret_cast<uint32_t*>(ctx->fwd_adj + ctx->e);
        const int64_t n_keys = one_sync ? UB : e_year;
        if (one_sync) launch_fill_tail_u32(s, keys, ctx->new_off + D, UB, (uint32_t)(N + D));
        launch_radix_sort_pairs(s, keys, reinterpret_cast<const uint32_t*>(ctx->src_of),
                                ctx->k_sorted, ctx->v_sorted, ctx->k_tmp, ctx->v_tmp, n_keys,
                                key_bits_for(one_sync ? N + D + 1 : N), ctx->scratch);
        const int nxt = 1 - ctx->cur;
        launch_bwd_offsets_sorted(s, N, D, ctx->bwd_off[ctx->cur], ctx->k_sorted, n_keys,
                                  ctx->bwd_off[nxt], ctx->in_deg);
        // the scratch holds at least max_papers + 2 int64 (its scan part)
        launch_bwd_merge(s, N, N + D, ctx->bwd_off[ctx->cur], ctx->bwd_adj[ctx->cur],
                         ctx->bwd_off[nxt], ctx->v_sorted, ctx->bwd_adj[nxt],
                         one_sync ? -1 : ctx->e + e_year, one_sync ? ctx->e + UB : 0,
                         ctx->scratch, ctx->cfg.max_papers + 2, ctx->tune_merge_out);
        pend(PG_K_CSR_APPEND);
        PG_HIP(hipGetLastError());
        ctx->cur = nxt;
        if (one_sync) {
            launch_acc_i64(s, ctx->new_off + D, sums + 4);
            PG_HIP(hipMemcpyAsync(ctx->h_sum, sums, 5 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
        }
    }
    PG_HIP(hipEventRecord(ctx->ev1, s));
    {
        int rc = seq_record_draws(ctx, e_year);
        if (rc) return rc;
    }
    PG_HIP(hipEventSynchronize(ctx->ev1));
    if (D > 0) {
        if (one_sync) {
            for (int k = 0; k < 4; ++k) hs[k] = ctx->h_sum[k];
            tot = (int64_t)ctx->h_sum[4];
        }
        e_year = tot;
        near_total = (int64_t)hs[0];
        attempts = (int64_t)hs[1];
        bfs_bytes = (double)hs[2];
    }
    float ms = 0.0f;
    PG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ctx->prof && D > 0) {
        const double lg = std::ceil(std::log2((double)std::max<int64_t>(N, 2)));
        const double e_old = (double)ctx->e;
        double bytes[PG_K_COUNT] = {};
        bytes[PG_K_WEIGHTS_CDF] = 20.0 * N;
        bytes[PG_K_PA_DRAW] = 8.0 * lg * attempts + 4.0 * D * (p->m - p->n_nb);
        bytes[PG_K_BFS] = bfs_bytes;
        bytes[PG_K_NB_DRAW] = 20.0 * D + 8.0 * D * p->m;
        bytes[PG_K_ER] = 4.0 * e_year + 4.0 * D;
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
        stats->near_boundary = near_total;
        stats->shortfall = (p->model == PG_MODEL_PA) ? D * (int64_t)p->m - e_year : 0;
        stats->pa_attempts = attempts;
        stats->device_ms = ms;
    }
    ctx->n = N + D;
    ctx->e += e_year;
    ctx->last_n_new = D;
    ctx->last_n_exist = N;
    ctx->last_m = p->model == PG_MODEL_PA ? p->m : 0;
    ctx->last_model = p->model;
    ctx->steps_done += 1;
    ctx->last_year = p->year;
    ctx->last_year_index = p->year_index;
    return PG_OK;
}

int pg_snapshot_save(pg_ctx* ctx) {
    if (!ctx) return PG_EINVAL;
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_snapshot_save: no graph loaded");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (!ctx->s_birth) {
        const int64_t N = ctx->cfg.max_papers, E = std::max<int64_t>(ctx->cfg.max_edges, 1);
        hipError_t r = hipSuccess;
        if (r == hipSuccess) r = dalloc(&ctx->s_birth, N);
        if (r == hipSuccess) r = dalloc(&ctx->s_in_deg, N);
        if (r == hipSuccess) r = dalloc(&ctx->s_fwd_off, N + 1);
        if (r == hipSuccess) r = dalloc(&ctx->s_fwd_adj, E);
        if (r == hipSuccess) r = dalloc(&ctx->s_bwd_off, N + 1);
        if (r == hipSuccess) r = dalloc(&ctx->s_bwd_adj, E);
        if (r != hipSuccess) return fail(ctx, PG_ENOMEM, "pg_snapshot_save: device allocation failed");
    }
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n, e = ctx->e;
    PG_HIP(hipMemcpyAsync(ctx->s_birth, ctx->birth, n * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->s_in_deg, ctx->in_deg, n * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->s_fwd_off, ctx->fwd_off, (n + 1) * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->s_bwd_off, ctx->bwd_off[ctx->cur], (n + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (e) {
        PG_HIP(hipMemcpyAsync(ctx->s_fwd_adj, ctx->fwd_adj, e * 4, hipMemcpyDeviceToDevice, s));
        PG_HIP(hipMemcpyAsync(ctx->s_bwd_adj, ctx->bwd_adj[ctx->cur], e * 4, hipMemcpyDeviceToDevice, s));
    }
    PG_HIP(hipStreamSynchronize(s));
    ctx->snap = true;
    ctx->snap_n = n;
    ctx->snap_e = e;
    ctx->snap_steps = ctx->steps_done;
    ctx->snap_last_year = ctx->last_year;
    ctx->snap_last_year_index = ctx->last_year_index;
    return PG_OK;
}

int pg_snapshot_restore(pg_ctx* ctx) {
    if (!ctx) return PG_EINVAL;
    if (!ctx->snap) return fail(ctx, PG_ESTATE, "pg_snapshot_restore: no snapshot saved");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->snap_n, e = ctx->snap_e;
    ctx->cur = 0;
    PG_HIP(hipMemcpyAsync(ctx->birth, ctx->s_birth, n * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->in_deg, ctx->s_in_deg, n * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->fwd_off, ctx->s_fwd_off, (n + 1) * 4, hipMemcpyDeviceToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->bwd_off[0], ctx->s_bwd_off, (n + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (e) {
        PG_HIP(hipMemcpyAsync(ctx->fwd_adj, ctx->s_fwd_adj, e * 4, hipMemcpyDeviceToDevice, s));
        PG_HIP(hipMemcpyAsync(ctx->bwd_adj[0], ctx->s_bwd_adj, e * 4, hipMemcpyDeviceToDevice, s));
    }
    PG_HIP(hipStreamSynchronize(s));
    ctx->n = n;
    ctx->e = e;
    ctx->last_model = -1;
    ctx->steps_done = ctx->snap_steps;
    ctx->last_year = ctx->snap_last_year;
    ctx->last_year_index = ctx->snap_last_year_index;
    return PG_OK;
}

int pg_profile_enable(pg_ctx* ctx, int on) {
    if (!ctx) return PG_EINVAL;
    ctx->prof = on != 0;
    return PG_OK;
}

int pg_profile_groups(pg_ctx* ctx, uint32_t mask) {
    if (!ctx) return PG_EINVAL;
    ctx->prof_mask = mask;
    return PG_OK;
}

int pg_profile_read(pg_ctx* ctx, int n, double* ms, double* bytes, int64_t* launches) {
    if (!ctx || n < 0) return PG_EINVAL;
    for (int k = 0; k < n && k < PG_K_COUNT; ++k) {
        if (ms) ms[k] = ctx->prof_ms[k];
        if (bytes) bytes[k] = ctx->prof_bytes[k];
        if (launches) launches[k] = ctx->prof_calls[k];
    }
    return PG_OK;
}

int pg_profile_reset(pg_ctx* ctx) {
    if (!ctx) return PG_EINVAL;
    for (int k = 0; k < PG_K_COUNT; ++k) {
        ctx->prof_ms[k] = 0.0;
        ctx->prof_bytes[k] = 0.0;
        ctx->prof_calls[k] = 0;
    }
    return PG_OK;
}

int pg_graph_size(const pg_ctx* ctx, int64_t* n_papers, int64_t* n_edges) {
    if (!ctx) return PG_EINVAL;
    if (n_papers) *n_papers = ctx->n;
    if (n_edges) *n_edges = ctx->e;
    return PG_OK;
}

int pg_graph_download(pg_ctx* ctx, int32_t* fwd_off, int32_t* fwd_adj, int32_t* bwd_off,
                      int32_t* bwd_adj, int32_t* in_deg, int32_t* birth, float* logfit) {
    if (!ctx) return PG_EINVAL;
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_graph_download: no graph loaded");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n, e = ctx->e;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    if (fwd_off) PG_HIP(cp(fwd_off, ctx->fwd_off, (n + 1) * 4));
    if (fwd_adj) PG_HIP(cp(fwd_adj, ctx->fwd_adj, e * 4));
    if (bwd_off) PG_HIP(cp(bwd_off, ctx->bwd_off[ctx->cur], (n + 1) * 4));
    if (bwd_adj) PG_HIP(cp(bwd_adj, ctx->bwd_adj[ctx->cur], e * 4));
    if (in_deg) PG_HIP(cp(in_deg, ctx->in_deg, n * 4));
    if (birth) PG_HIP(cp(birth, ctx->birth, n * 4));
    if (logfit) PG_HIP(cp(logfit, ctx->logfit, n * 4));
    PG_HIP(hipStreamSynchronize(s));
    return PG_OK;
}

int pg_year_citations(pg_ctx* ctx, int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* near_boundary) {
    if (!ctx) return PG_EINVAL;
    if (ctx->last_model != PG_MODEL_PA)
        return fail(ctx, PG_ESTATE, "pg_year_citations: the last step was not a PA year");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    const int64_t D = ctx->last_n_new, M = ctx->last_m;
    if (cit && D * M) PG_HIP(hipMemcpyAsync(cit, ctx->cit, D * M * 4, hipMemcpyDeviceToHost, s));
    if (kind && D * M) PG_HIP(hipMemcpyAsync(kind, ctx->kind, D * M, hipMemcpyDeviceToHost, s));
    if (cnt && D) PG_HIP(hipMemcpyAsync(cnt, ctx->cnt, D * 4, hipMemcpyDeviceToHost, s));
    if (near_boundary && D) PG_HIP(hipMemcpyAsync(near_boundary, ctx->near, D * 4, hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    return PG_OK;
}

int pg_year_weights(pg_ctx* ctx, float* w, double* cdf) {
    if (!ctx) return PG_EINVAL;
    if (ctx->last_model != PG_MODEL_PA)
        return fail(ctx, PG_ESTATE, "pg_year_weights: the last step was not a PA year");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    hipStream_t s = ctx->stream;
    const int64_t N = ctx->last_n_exist;
    if (w) PG_HIP(hipMemcpyAsync(w, ctx->w, N * 4, hipMemcpyDeviceToHost, s));
    if (cdf) PG_HIP(hipMemcpyAsync(cdf, ctx->cdf, N * 8, hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    return PG_OK;
}

int pg_set_tuning(pg_ctx* ctx, const pg_tuning* t) {
    if (!ctx) return PG_EINVAL;
    if (!t || t->size != sizeof(pg_tuning))
        return fail(ctx, PG_EINVAL, "pg_set_tuning: null or struct size mismatch");
    if (const char* e = tuning_error(t)) return fail(ctx, PG_EINVAL, std::string("pg_set_tuning: ") + e);
    if (t->lik_lanes != ctx->tune_lik_lanes) ctx->seq.tables_dirty = true;
    ctx->tune_pa_lanes = t->pa_lanes;
    ctx->tune_lik_lanes = t->lik_lanes;
    ctx->tune_bfs_mode = t->bfs_mode;
    ctx->tune_small_max = t->bfs_small_max;
    ctx->tune_mid_max = t->bfs_mid_max;
    ctx->tune_merge_out = t->merge_out;
    return PG_OK;
}

int pg_year_autotune(pg_ctx* ctx, const pg_year_params* years, int32_t n_years, int32_t reps,
                     pg_tune_report* report) {
    if (!ctx) return PG_EINVAL;
    if (report && report->size != sizeof(pg_tune_report))
        return fail(ctx, PG_EINVAL, "pg_year_autotune: pg_tune_report.size mismatch");
    if (!years || n_years < 1 || reps < 1 || reps > 64)
        return fail(ctx, PG_EINVAL, "pg_year_autotune: need years and reps in [1, 64]");
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_year_autotune: no graph loaded");
    if (ctx->seq_recording) return fail(ctx, PG_ESTATE, "pg_year_autotune: sequence recording is on");
    if (int rc = pg_snapshot_save(ctx)) return rc;
    const pg_tuning before{sizeof(pg_tuning), ctx->tune_pa_lanes, ctx->tune_lik_lanes,
                           ctx->tune_bfs_mode, ctx->tune_small_max, ctx->tune_mid_max,
                           ctx->tune_merge_out};
    std::vector<pg_tuning> cand;
    std::vector<double> ms;
    int64_t ref[3] = {0, 0, 0};
    bool have_ref = false;
    // one candidate: an untimed pass over the years, then reps timed ones,
    // each from the saved state
    auto trial = [&](const pg_tuning& t) -> int {
        if (int rc = pg_set_tuning(ctx, &t)) return rc;
        std::vector<double> v;
        for (int r = 0; r <= reps; ++r) {
            if (int rc = pg_snapshot_restore(ctx)) return rc;
            int64_t got[3] = {0, 0, 0};
            double dms = 0.0;
            for (int32_t y = 0; y < n_years; ++y) {
                pg_year_stats st{};
                st.size = sizeof st;
                if (int rc = pg_year_step(ctx, &years[y], &st)) return rc;
                got[0] += st.edges_added;
                got[1] += st.near_boundary;
                got[2] += st.pa_attempts;
                dms += st.device_ms;
            }
            if (!have_ref) {
                for (int k = 0; k < 3; ++k) ref[k] = got[k];
                have_ref = true;
            }
            if (got[0] != ref[0] || got[1] != ref[1] || got[2] != ref[2])
                return fail(ctx, PG_ESTATE, "pg_year_autotune: a launch shape changed the run");
            if (r > 0) v.push_back(dms);
        }
        std::sort(v.begin(), v.end());
        cand.push_back(t);
        ms.push_back(v.size() % 2 ? v[v.size() / 2] : 0.5 * (v[v.size() / 2 - 1] + v[v.size() / 2]));
        return PG_OK;
    };
    auto run_all = [&]() -> int {
        int m = 0;
        for (int32_t y = 0; y < n_years; ++y) m = std::max(m, (int)years[y].m - (int)years[y].n_nb);
        const int def_lanes = pa_draw_lanes(m);
        double default_ms = 0.0;
        size_t best = 0;
        for (int l : {8, 16, 32, 64}) {
            pg_tuning t = before;
            t.pa_lanes = l;
            t.bfs_mode = 0;
            t.bfs_small_max = t.bfs_mid_max = t.merge_out = 0;
            if (int rc = trial(t)) return rc;
            if (l == def_lanes) default_ms = ms.back();
        }
        auto pick = [&]() { best = std::min_element(ms.begin(), ms.end()) - ms.begin(); };
        pick();
        auto vary = [&](int which, std::initializer_list<int> vals) -> int {
            for (int v : vals) {
                pg_tuning t = cand[best];
                (which == 0 ? t.bfs_mode : which == 1 ? t.bfs_small_max
                                         : which == 2 ? t.bfs_mid_max : t.merge_out) = v;
                if (int rc = trial(t)) return rc;
            }
            pick();
            return PG_OK;
        };
        if (int rc = vary(0, {1, 2})) return rc;
        if (int rc = vary(1, {512, 1024})) return rc;
        if (int rc = vary(2, {2048})) return rc;
        if (int rc = vary(3, {512, 2048})) return rc;
        const pg_tuning chosen = cand[best];
        if (int rc = pg_set_tuning(ctx, &chosen)) return rc;
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
    };
    const int rc = run_all();
    if (rc) (void)pg_set_tuning(ctx, &before);
    // back to the state of the call (kept on failure too, when restorable)
    const int rr = pg_snapshot_restore(ctx);
    return rc ? rc : rr;
}

int pg_bfs_timed(pg_ctx* ctx, int64_t count, const int32_t* roots, int32_t* out, int32_t* out_len,
                 int64_t* ticks) {
    if (!ctx) return PG_EINVAL;
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_bfs: no graph loaded");
    if (count < 0 || (count > 0 && (!roots || !out || !out_len)))
        return fail(ctx, PG_EINVAL, "pg_bfs: bad arguments");
    if (count == 0) return PG_OK;
    for (int64_t i = 0; i < count; ++i)
        if (roots[i] < 0 || roots[i] >= ctx->n) return fail(ctx, PG_EINVAL, "pg_bfs: root out of range");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (int rcs = ensure_bfs_stage(ctx)) return rcs;
    hipStream_t s = ctx->stream;
    const int cap = ctx->cfg.bfs_cap;
    int32_t *d_roots = nullptr, *d_out = nullptr, *d_len = nullptr;
    int64_t* d_ticks = nullptr;
    hipError_t rc = dalloc(&d_roots, count);
    if (rc == hipSuccess) rc = dalloc(&d_out, count * cap);
    if (rc == hipSuccess) rc = dalloc(&d_len, count);
    if (rc == hipSuccess && ticks) rc = dalloc(&d_ticks, 3 * count);
    if (rc != hipSuccess) {
        (void)hipFree(d_roots); (void)hipFree(d_out); (void)hipFree(d_len); (void)hipFree(d_ticks);
        return fail(ctx, PG_ENOMEM, "pg_bfs: device allocation failed");
    }
    rc = hipMemcpyAsync(d_roots, roots, count * 4, hipMemcpyHostToDevice, s);
    if (rc == hipSuccess) {
        launch_bfs_only(s, view(ctx), count, d_roots, cap, d_out, d_len, d_ticks,
                        bfs_stage_view(ctx->stage, ctx->tune_bfs_mode));
        rc = hipGetLastError();
    }
    if (rc == hipSuccess) rc = hipMemcpyAsync(out, d_out, count * cap * 4, hipMemcpyDeviceToHost, s);
    if (rc == hipSuccess) rc = hipMemcpyAsync(out_len, d_len, count * 4, hipMemcpyDeviceToHost, s);
    if (rc == hipSuccess && ticks)
        rc = hipMemcpyAsync(ticks, d_ticks, 3 * count * 8, hipMemcpyDeviceToHost, s);
    if (rc == hipSuccess) rc = hipStreamSynchronize(s);
    (void)hipFree(d_roots); (void)hipFree(d_out); (void)hipFree(d_len); (void)hipFree(d_ticks);
    if (rc != hipSuccess) return fail(ctx, PG_EHIP, std::string("pg_bfs: ") + hipGetErrorString(rc));
    return PG_OK;
}

int pg_bfs(pg_ctx* ctx, int64_t count, const int32_t* roots, int32_t* out, int32_t* out_len) {
    return pg_bfs_timed(ctx, count, roots, out, out_len, nullptr);
}

}  // extern "C"

const char* tuning_error(const pg_tuning* t) {
    auto in = [](int v, std::initializer_list<int> ok) {
        for (int x : ok)
            if (v == x) return true;
        return false;
    };
    if (!in(t->pa_lanes, {0, 8, 16, 32, 64})) return "pa_lanes must be 0, 8, 16, 32 or 64";
    if (!in(t->lik_lanes, {0, 1, 2, 4, 8})) return "lik_lanes must be 0, 1, 2, 4 or 8";
    if (t->bfs_mode < 0 || t->bfs_mode > 2) return "bfs_mode must be 0, 1 or 2";
    if (!in(t->bfs_small_max, {0, 256, 512, 1024, 2048}))
        return "bfs_small_max must be 0, 256, 512, 1024 or 2048";
    if (!in(t->bfs_mid_max, {0, 512, 1024, 2048, 4096}))
        return "bfs_mid_max must be 0, 512, 1024, 2048 or 4096";
    if (!in(t->merge_out, {0, 256, 512, 1024, 2048}))
        return "merge_out must be 0, 256, 512, 1024 or 2048";
    return nullptr;
}

extern "C" {

int64_t pg_growth_count(int64_t n, double rate) {
    return (int64_t)std::ceil((double)n * rate);
}

}  // extern "C"
