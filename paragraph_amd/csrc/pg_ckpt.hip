// pg_ckpt.hip -- exact year-granular checkpoint and resume of a context's
// graph state through a file, and the progress counters.
//
// No reference counterpart: the reference rebuilds everything per run
// (src/kernel.cu:1682-1845) and keeps nothing between runs.  Draws are
// keyed by (seed, paper, attempt, year index) in a counter-based Philox
// stream, so the state below is all a later year depends on: a context that
// loads a checkpoint continues bit for bit like the one that wrote it.
//
// File layout (host byte order): a fixed header, then birth[n], in_deg[n],
// fwd_off[n+1], fwd_adj[e], bwd_off[n+1], bwd_adj[e] (int32), logfit
// [n_logfit] (float) and the caller's blob; the header carries an FNV-1a
// 64-bit checksum of everything after it.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paragraph.h"
#include "pg_ctx.h"
#include "pg_internal.h"

namespace {

constexpr char kMagic[8] = {'P', 'G', 'C', 'K', 'P', 'T', '0', '1'};
constexpr uint32_t kVersion = 1;

struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint64_t seed;
    int64_t n, e, n_logfit, user_bytes;
    int32_t steps_done, last_year, last_year_index, reserved;
    uint64_t checksum;
};

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t k = 0; k < n; ++k) {
            h ^= c[k];
            h *= 0x100000001b3ull;
        }
    }
};

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

bool put(std::FILE* f, Fnv* sum, const void* p, size_t n) {
    if (sum) sum->add(p, n);
    return n == 0 || std::fwrite(p, 1, n, f) == n;
}

bool get(std::FILE* f, Fnv* sum, void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n) return false;
    if (sum) sum->add(p, n);
    return true;
}

}  // namespace

extern "C" {

int pg_get_progress(pg_ctx* ctx, int32_t* steps_done, int32_t* last_year,
                    int32_t* last_year_index) {
    if (!ctx) return PG_EINVAL;
    if (steps_done) *steps_done = ctx->steps_done;
    if (last_year) *last_year = ctx->last_year;
    if (last_year_index) *last_year_index = ctx->last_year_index;
    return PG_OK;
}

int pg_checkpoint_save(pg_ctx* ctx, const char* path, const void* user, int64_t user_bytes) {
    if (!ctx) return PG_EINVAL;
    if (!path || user_bytes < 0 || (user_bytes > 0 && !user))
        return fail(ctx, PG_EINVAL, "pg_checkpoint_save: bad arguments");
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_checkpoint_save: no graph loaded");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    const int64_t n = ctx->n, e = ctx->e, nl = ctx->n_logfit;
    std::vector<int32_t> birth(n), indeg(n), fo(n + 1), fa(e), bo(n + 1), ba(e);
    std::vector<float> lf(nl);
    hipStream_t s = ctx->stream;
    PG_HIP(hipMemcpyAsync(birth.data(), ctx->birth, n * 4, hipMemcpyDeviceToHost, s));
    PG_HIP(hipMemcpyAsync(indeg.data(), ctx->in_deg, n * 4, hipMemcpyDeviceToHost, s));
    PG_HIP(hipMemcpyAsync(fo.data(), ctx->fwd_off, (n + 1) * 4, hipMemcpyDeviceToHost, s));
    PG_HIP(hipMemcpyAsync(bo.data(), ctx->bwd_off[ctx->cur], (n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (e) {
        PG_HIP(hipMemcpyAsync(fa.data(), ctx->fwd_adj, e * 4, hipMemcpyDeviceToHost, s));
        PG_HIP(hipMemcpyAsync(ba.data(), ctx->bwd_adj[ctx->cur], e * 4, hipMemcpyDeviceToHost, s));
    }
    if (nl) PG_HIP(hipMemcpyAsync(lf.data(), ctx->logfit, nl * 4, hipMemcpyDeviceToHost, s));
    PG_HIP(hipStreamSynchronize(s));
    CkptHeader h{};
    std::memcpy(h.magic, kMagic, sizeof kMagic);
    h.version = kVersion;
    h.header_bytes = sizeof h;
    h.seed = ctx->cfg.seed;
    h.n = n;
    h.e = e;
    h.n_logfit = nl;
    h.user_bytes = user_bytes;
    h.steps_done = ctx->steps_done;
    h.last_year = ctx->last_year;
    h.last_year_index = ctx->last_year_index;
    Fnv sum;
    sum.add(birth.data(), n * 4);
    sum.add(indeg.data(), n * 4);
    sum.add(fo.data(), (n + 1) * 4);
    sum.add(fa.data(), e * 4);
    sum.add(bo.data(), (n + 1) * 4);
    sum.add(ba.data(), e * 4);
    sum.add(lf.data(), nl * 4);
    sum.add(user, (size_t)user_bytes);
    h.checksum = sum.h;
    const std::string tmp = std::string(path) + ".tmp";
    {
        File f;
        f.f = std::fopen(tmp.c_str(), "wb");
        if (!f.f) return fail(ctx, PG_EIO, "pg_checkpoint_save: cannot open " + tmp);
        bool ok = put(f.f, nullptr, &h, sizeof h) && put(f.f, nullptr, birth.data(), n * 4) &&
                  put(f.f, nullptr, indeg.data(), n * 4) &&
                  put(f.f, nullptr, fo.data(), (n + 1) * 4) && put(f.f, nullptr, fa.data(), e * 4) &&
                  put(f.f, nullptr, bo.data(), (n + 1) * 4) && put(f.f, nullptr, ba.data(), e * 4) &&
                  put(f.f, nullptr, lf.data(), nl * 4) && put(f.f, nullptr, user, (size_t)user_bytes);
        ok = ok && std::fflush(f.f) == 0;
        ok = std::fclose(f.f) == 0 && ok;
        f.f = nullptr;
        if (!ok) {
            std::remove(tmp.c_str());
            return fail(ctx, PG_EIO, "pg_checkpoint_save: write to " + tmp + " failed");
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(ctx, PG_EIO, std::string("pg_checkpoint_save: cannot rename to ") + path);
    }
    return PG_OK;
}

// The checksum only catches accidental damage, so the state is checked for
// what the kernels rely on before any of it reaches the device (O(n + e)):
// offsets monotone from 0 to e, ids in [0, n), rows strictly ascending and
// free of self-citations, the backward CSR the transpose of the forward one
// (citers ascending), and in-degrees equal to the backward row lengths.
// Returns null when consistent, else what failed.
static const char* check_state(int64_t n, int64_t e, const std::vector<int32_t>& birth,
                               const std::vector<int32_t>& indeg, const std::vector<int32_t>& fo,
                               const std::vector<int32_t>& fa, const std::vector<int32_t>& bo,
                               const std::vector<int32_t>& ba) {
    (void)birth;
    if (fo[0] != 0 || fo[n] != e || bo[0] != 0 || bo[n] != e) return "offset ends";
    for (int64_t v = 0; v < n; ++v) {
        if (fo[v + 1] < fo[v] || bo[v + 1] < bo[v]) return "offsets not monotone";
        if (indeg[v] != bo[v + 1] - bo[v]) return "in-degree is not the backward row length";
    }
    // forward rows: ids in range, strictly ascending, no self-citation; count
    // each target's citers for the transpose check
    std::vector<int32_t> cur(bo.begin(), bo.end() - 1);
    for (int64_t v = 0; v < n; ++v) {
        for (int64_t k = fo[v]; k < fo[v + 1]; ++k) {
            const int32_t t = fa[k];
            if (t < 0 || t >= n) return "forward id out of range";
            if (t == v) return "self-citation";
            if (k > fo[v] && fa[k - 1] >= t) return "forward row not strictly ascending";
            // citers of t arrive in ascending v, so they must fill t's
            // backward row in order
            if (cur[t] >= bo[t + 1] || ba[cur[t]] != (int32_t)v)
                return "backward rows are not the transpose of the forward rows";
            ++cur[t];
        }
    }
    for (int64_t t = 0; t < n; ++t)
        if (cur[t] != bo[t + 1]) return "backward rows are not the transpose of the forward rows";
    return nullptr;
}

int pg_checkpoint_load(pg_ctx* ctx, const char* path, void* user, int64_t user_cap,
                       int64_t* user_bytes) {
    if (!ctx) return PG_EINVAL;
    if (!path || user_cap < 0 || (user_cap > 0 && !user))
        return fail(ctx, PG_EINVAL, "pg_checkpoint_load: bad arguments");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return fail(ctx, PG_EIO, std::string("pg_checkpoint_load: cannot open ") + path);
    CkptHeader h{};
    if (!get(f.f, nullptr, &h, sizeof h) || std::memcmp(h.magic, kMagic, sizeof kMagic) != 0 ||
        h.version != kVersion || h.header_bytes != sizeof h)
        return fail(ctx, PG_EIO, std::string("pg_checkpoint_load: not a version-1 checkpoint: ") + path);
    if (h.seed != ctx->cfg.seed)
        return fail(ctx, PG_EINVAL, "pg_checkpoint_load: checkpoint written with Philox seed " +
                                        std::to_string(h.seed) + ", context has " +
                                        std::to_string(ctx->cfg.seed));
    if (h.n < 1 || h.e < 0 || h.n_logfit < h.n || h.user_bytes < 0)
        return fail(ctx, PG_EIO, "pg_checkpoint_load: malformed header");
    if (h.n > ctx->cfg.max_papers || h.e > ctx->cfg.max_edges || h.n_logfit > ctx->cfg.max_papers)
        return fail(ctx, PG_ECAP, "pg_checkpoint_load: the state exceeds this context's capacities");
    if (h.user_bytes > user_cap)
        return fail(ctx, PG_ECAP, "pg_checkpoint_load: the caller blob exceeds user_cap");
    const int64_t n = h.n, e = h.e, nl = h.n_logfit;
    std::vector<int32_t> birth(n), indeg(n), fo(n + 1), fa(e), bo(n + 1), ba(e);
    std::vector<float> lf(nl);
    std::vector<char> ub((size_t)h.user_bytes);
    Fnv sum;
    const bool ok = get(f.f, &sum, birth.data(), n * 4) && get(f.f, &sum, indeg.data(), n * 4) &&
                    get(f.f, &sum, fo.data(), (n + 1) * 4) && get(f.f, &sum, fa.data(), e * 4) &&
                    get(f.f, &sum, bo.data(), (n + 1) * 4) && get(f.f, &sum, ba.data(), e * 4) &&
                    get(f.f, &sum, lf.data(), nl * 4) && get(f.f, &sum, ub.data(), ub.size());
    if (!ok) return fail(ctx, PG_EIO, std::string("pg_checkpoint_load: truncated file ") + path);
    if (sum.h != h.checksum) return fail(ctx, PG_EIO, "pg_checkpoint_load: checksum mismatch");
    if (const char* why = check_state(n, e, birth, indeg, fo, fa, bo, ba))
        return fail(ctx, PG_EIO, std::string("pg_checkpoint_load: inconsistent graph state (") +
                                     why + ")");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (!ctx->logfit) {
        PG_HIP(dalloc(&ctx->logfit, ctx->cfg.max_papers));
        ctx->logfit_owned = false;
    }
    hipStream_t s = ctx->stream;
    ctx->loaded = false;  // until the copies below complete
    ctx->cur = 0;
    PG_HIP(hipMemcpyAsync(ctx->birth, birth.data(), n * 4, hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->in_deg, indeg.data(), n * 4, hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->fwd_off, fo.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
    PG_HIP(hipMemcpyAsync(ctx->bwd_off[0], bo.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
    if (e) {
        PG_HIP(hipMemcpyAsync(ctx->fwd_adj, fa.data(), e * 4, hipMemcpyHostToDevice, s));
        PG_HIP(hipMemcpyAsync(ctx->bwd_adj[0], ba.data(), e * 4, hipMemcpyHostToDevice, s));
    }
    PG_HIP(hipMemcpyAsync(ctx->logfit, lf.data(), nl * 4, hipMemcpyHostToDevice, s));
    PG_HIP(hipStreamSynchronize(s));
    ctx->n = n;
    ctx->e = e;
    // log-fitness of papers beyond the checkpoint's (supplied to this
    // context by pg_graph_load for a longer run) is kept
    ctx->n_logfit = std::max(nl, ctx->logfit_owned ? ctx->n_logfit : int64_t(0));
    ctx->logfit_owned = true;
    ctx->loaded = true;
    ctx->last_model = -1;
    ctx->steps_done = h.steps_done;
    ctx->last_year = h.last_year;
    ctx->last_year_index = h.last_year_index;
    if (user && !ub.empty()) std::memcpy(user, ub.data(), ub.size());
    if (user_bytes) *user_bytes = h.user_bytes;
    ctx->err.clear();
    return PG_OK;
}

}  // extern "C"
