// pg_lik.hip -- host side of the citation-sequence store and the PA
// likelihood entry points declared in include/paragraph.h (no reference
// counterpart).  Device layout of the store (one per context):
//   deg   : int32[paper-years]  in-degree snapshots, year after year
//   cit   : int32[new papers * m], kind : int8[new papers * m]
//   cnt   : int32[new papers],  birth : int32[papers]
//   per-year tables n_exist/n_new/deg_off/new_off (int64), year_of (int32)
//   slot-pass block table (year, first paper) and per-year block ranges
//   partials: Y*nP (paper pass) + 4*nE (slot pass) + 2*nP (backward pass)
//   doubles, per-year sums [Y][5], the scalar result [3].
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

void free_ptr(void* p) {
    if (p) (void)hipFree(p);
}

void seq_free_batch(pg_seq_store& q) {
    free_ptr(q.b_ev_block); free_ptr(q.b_year_blk); free_ptr(q.b_part); free_ptr(q.b_year_sums);
    free_ptr(q.b_share); free_ptr(q.b_evt); free_ptr(q.b_at); free_ptr(q.b_out);
    free_ptr(q.b_job_start); free_ptr(q.b_job_tag); free_ptr(q.b_bigpart);
    free_ptr(q.b_big_paper); free_ptr(q.b_big_first);
    q.b_ev_block = q.b_year_blk = nullptr;
    q.b_part = q.b_year_sums = q.b_share = q.b_evt = q.b_at = q.b_out = nullptr;
    q.b_job_start = nullptr;
    q.b_job_tag = q.b_big_paper = q.b_big_first = nullptr;
    q.b_bigpart = nullptr;
    q.b_n_jobs = 0;
    q.b_n_big = 0;
    q.bB = 0;
    for (int k = 0; k < 4; ++k) q.b_key[k] = -1;
}

void seq_free_tables(pg_seq_store& q) {
    seq_free_batch(q);
    free_ptr(q.d_n_exist); free_ptr(q.d_n_new); free_ptr(q.d_deg_off); free_ptr(q.d_new_off);
    free_ptr(q.d_year_of); free_ptr(q.d_ev_block); free_ptr(q.d_ev_year_blk);
    free_ptr(q.part); free_ptr(q.year_sums); free_ptr(q.d_scalars);
    free_ptr(q.skey); free_ptr(q.pos_t); free_ptr(q.cit_t); free_ptr(q.kind_t);
    free_ptr(q.seg_off); free_ptr(q.share); free_ptr(q.evt); free_ptr(q.carry);
    free_ptr(q.la_t); free_ptr(q.Lpy);
    q.skey = q.pos_t = nullptr;
    q.cit_t = q.seg_off = nullptr;
    q.kind_t = nullptr;
    q.la_t = nullptr;
    q.Lpy = nullptr;
    q.share = q.evt = q.carry = nullptr;
    q.carry_cap = 0;
    q.d_n_exist = q.d_n_new = q.d_deg_off = q.d_new_off = nullptr;
    q.d_year_of = q.d_ev_block = q.d_ev_year_blk = nullptr;
    q.part = q.year_sums = q.d_scalars = nullptr;
    q.part_cap = q.table_cap_years = q.table_cap_blocks = 0;
    q.tables_dirty = true;
}

void seq_clear(pg_seq_store& q) {
    q.n_exist.clear(); q.n_new.clear(); q.deg_off.clear(); q.new_off.clear(); q.year_of.clear();
    q.n_papers = q.paper_years = q.new_total = q.filled = 0;
    q.pending = false;
    q.tables_dirty = true;
}

LikSeq seq_view(const pg_seq_store& q);

// Per-year tables, slot-pass block table and partial buffers on the device.
int seq_prepare(pg_ctx* ctx) {
    pg_seq_store& q = ctx->seq;
    if (!q.tables_dirty) return PG_OK;
    const int64_t Y = (int64_t)q.n_exist.size();
    const int64_t nP = lik_paper_blocks(q.n_papers);
    const LikCiteShape sh = lik_cite_shape(q.m, ctx->tune_lik_lanes);
    const int64_t per_block = 256 / sh.G;
    std::vector<int32_t> ev_block, year_blk(Y + 1, 0);
    for (int64_t y = 0; y < Y; ++y) {
        year_blk[y] = (int32_t)(ev_block.size() / 2);
        for (int64_t f = 0; f < q.n_new[y]; f += per_block) {
            ev_block.push_back((int32_t)y);
            ev_block.push_back((int32_t)f);
        }
    }
    year_blk[Y] = (int32_t)(ev_block.size() / 2);
    const int64_t nE = (int64_t)ev_block.size() / 2;
    const int64_t part_need = Y * nP + 4 * nE + 2 * nP + 1;
    if (Y + 1 > q.table_cap_years || nE > q.table_cap_blocks || part_need > q.part_cap) {
        const int64_t ycap = std::max<int64_t>(Y + 1, 2 * q.table_cap_years);
        const int64_t bcap = std::max<int64_t>(std::max<int64_t>(nE, 1), 2 * q.table_cap_blocks);
        const int64_t pcap = std::max<int64_t>(part_need, 2 * q.part_cap);
        seq_free_tables(q);
        hipError_t r = dalloc(&q.d_n_exist, ycap);
        if (r == hipSuccess) r = dalloc(&q.d_n_new, ycap);
        if (r == hipSuccess) r = dalloc(&q.d_deg_off, ycap);
        if (r == hipSuccess) r = dalloc(&q.d_new_off, ycap);
        if (r == hipSuccess) r = dalloc(&q.d_year_of, ycap);
        if (r == hipSuccess) r = dalloc(&q.d_ev_year_blk, ycap + 1);
        if (r == hipSuccess) r = dalloc(&q.d_ev_block, 2 * bcap);
        if (r == hipSuccess) r = dalloc(&q.part, pcap);
        if (r == hipSuccess) r = dalloc(&q.year_sums, ycap * kLikYearCols);
        if (r == hipSuccess) r = dalloc(&q.d_scalars, 4);
        if (r != hipSuccess) {
            seq_free_tables(q);
            return fail(ctx, PG_ENOMEM, "likelihood tables: device allocation failed");
        }
        q.table_cap_years = ycap;
        q.table_cap_blocks = bcap;
        q.part_cap = pcap;
    }
    if (Y > 0) {
        PG_HIP(hipMemcpy(q.d_n_exist, q.n_exist.data(), Y * 8, hipMemcpyHostToDevice));
        PG_HIP(hipMemcpy(q.d_n_new, q.n_new.data(), Y * 8, hipMemcpyHostToDevice));
        PG_HIP(hipMemcpy(q.d_deg_off, q.deg_off.data(), Y * 8, hipMemcpyHostToDevice));
        PG_HIP(hipMemcpy(q.d_new_off, q.new_off.data(), Y * 8, hipMemcpyHostToDevice));
        PG_HIP(hipMemcpy(q.d_year_of, q.year_of.data(), Y * 4, hipMemcpyHostToDevice));
    }
    PG_HIP(hipMemcpy(q.d_ev_year_blk, year_blk.data(), (Y + 1) * 4, hipMemcpyHostToDevice));
    if (nE > 0)
        PG_HIP(hipMemcpy(q.d_ev_block, ev_block.data(), nE * 8, hipMemcpyHostToDevice));
    q.n_ev_blocks = nE;
    q.n_paper_blocks = nP;
    q.cite_shape = sh;
    // slots sorted by target, for the per-paper sums of the dL/dphi shares
    const int64_t T = q.new_total * q.m;
    free_ptr(q.skey); free_ptr(q.pos_t); free_ptr(q.cit_t); free_ptr(q.kind_t);
    free_ptr(q.seg_off); free_ptr(q.share); free_ptr(q.evt); free_ptr(q.carry);
    free_ptr(q.la_t); free_ptr(q.Lpy);
    q.skey = q.pos_t = nullptr;
    q.cit_t = q.seg_off = nullptr;
    q.kind_t = nullptr;
    q.la_t = nullptr;
    q.Lpy = nullptr;
    q.share = q.evt = q.carry = nullptr;
    const int64_t nchunk = lik_seg_chunks(q.filled);
    uint32_t *keys = nullptr, *vals = nullptr, *k_tmp = nullptr, *v_tmp = nullptr,
             *perm = nullptr;
    int64_t* scratch = nullptr;
    hipError_t r = dalloc(&q.skey, T);
    if (r == hipSuccess) r = dalloc(&q.pos_t, T);
    if (r == hipSuccess) r = dalloc(&q.cit_t, T);
    if (r == hipSuccess) r = dalloc(&q.kind_t, T);
    if (r == hipSuccess) r = dalloc(&q.la_t, T);
    if (r == hipSuccess) r = dalloc(&q.Lpy, std::max<int64_t>(q.paper_years, 1));
    if (r == hipSuccess) r = dalloc(&q.seg_off, q.n_papers + 1);
    if (r == hipSuccess) r = dalloc(&q.share, q.filled);
    if (r == hipSuccess) r = dalloc(&perm, T);
    if (r == hipSuccess) r = dalloc(&q.evt, q.n_papers);
    if (r == hipSuccess) r = dalloc(&q.carry, 2 * nchunk);
    if (r == hipSuccess) r = dalloc(&keys, T);
    if (r == hipSuccess) r = dalloc(&vals, T);
    if (r == hipSuccess) r = dalloc(&k_tmp, T);
    if (r == hipSuccess) r = dalloc(&v_tmp, T);
    if (r == hipSuccess) r = dalloc(&scratch, radix_scratch_elems(T));
    if (r == hipSuccess) {
        int bits = 1;
        while (bits < 31 && (int64_t(1) << bits) <= q.n_papers) ++bits;
        launch_lik_sort_slots(ctx->stream, T, q.m, q.n_papers, q.cnt, q.cit, q.kind, keys, vals,
                              q.skey, perm, q.cit_t, q.kind_t, q.pos_t, k_tmp, v_tmp, bits,
                              scratch, q.seg_off);
        r = hipGetLastError();
        if (r == hipSuccess) {
            launch_lik_slot_la(ctx->stream, seq_view(q), q.la_t);
            launch_lik_deg_log(ctx->stream, q.deg, q.paper_years, q.Lpy);
            r = hipGetLastError();
        }
        if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    }
    free_ptr(keys); free_ptr(vals); free_ptr(k_tmp); free_ptr(v_tmp); free_ptr(scratch);
    free_ptr(perm);
    if (r != hipSuccess) {
        seq_free_tables(q);
        return fail(ctx, PG_EHIP, std::string("likelihood slot sort: ") + hipGetErrorString(r));
    }
    q.carry_cap = 2 * nchunk;
    q.tables_dirty = false;
    return PG_OK;
}

LikSeq seq_view(const pg_seq_store& q) {
    LikSeq v;
    v.n_years = (int32_t)q.n_exist.size();
    v.m = q.m;
    v.n_papers = q.n_papers;
    v.n_exist = q.d_n_exist;
    v.n_new = q.d_n_new;
    v.deg_off = q.d_deg_off;
    v.new_off = q.d_new_off;
    v.year_of = q.d_year_of;
    v.deg = q.deg;
    v.cit = q.cit;
    v.kind = q.kind;
    v.cnt = q.cnt;
    v.birth = q.birth;
    v.n_cite = q.new_total;
    v.cit_t = q.cit_t;
    v.kind_t = q.kind_t;
    v.pos_t = q.pos_t;
    v.la_t = q.la_t;
    v.Lpy = q.Lpy;
    return v;
}

// Checks a new year against the store's capacities.
int seq_check_year(pg_ctx* ctx, const char* fn, int64_t n_exist, int64_t n_new) {
    const pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, std::string(fn) + ": no sequence reserved");
    if (n_exist < 0 || n_new < 0) return fail(ctx, PG_EINVAL, std::string(fn) + ": negative size");
    if ((int64_t)q.n_exist.size() + 1 > kLikMaxYears)
        return fail(ctx, PG_ECAP, std::string(fn) + ": more than 256 years");
    if (q.paper_years + n_exist > q.cap_paper_years || q.new_total + n_new > q.cap_new ||
        n_exist + n_new > q.cap_papers)
        return fail(ctx, PG_ECAP, std::string(fn) + ": sequence capacity exceeded");
    return PG_OK;
}

void seq_commit_year(pg_seq_store& q, int32_t year, int64_t n_exist, int64_t n_new) {
    q.n_exist.push_back(n_exist);
    q.n_new.push_back(n_new);
    q.deg_off.push_back(q.paper_years);
    q.new_off.push_back(q.new_total);
    q.year_of.push_back(year);
    q.paper_years += n_exist;
    q.new_total += n_new;
    q.tables_dirty = true;
}

// Algorithmic HBM bytes of one evaluation (the formulas in paragraph.h).
double lik_bytes(const pg_seq_store& q, bool grad) {
    const double PY = (double)q.paper_years, P = (double)q.n_papers, NP = (double)q.new_total,
                 S = (double)q.filled;
    return grad ? 8.0 * PY + 44.0 * P + 4.0 * NP + 33.0 * S
                : 4.0 * PY + 8.0 * P + 4.0 * NP + 17.0 * S;
}

// Buffers and the slot-pass block table of the batched evaluation at B
// points for the current sequence (seq_prepare done).
int batch_prepare(pg_ctx* ctx, int B) {
    pg_seq_store& q = ctx->seq;
    const int64_t Y = (int64_t)q.n_exist.size();
    const int64_t key[4] = {Y, q.new_total, q.filled, q.n_papers};
    if (q.bB == B && std::equal(key, key + 4, q.b_key)) return PG_OK;
    seq_free_batch(q);
    const LikCiteShape sh = lik_batch_cite_shape(q.m, B);
    const int64_t per_block = kLikBCitePapers;
    std::vector<int32_t> ev_block, year_blk(Y + 1, 0);
    for (int64_t y = 0; y < Y; ++y) {
        year_blk[y] = (int32_t)(ev_block.size() / 2);
        for (int64_t f = 0; f < q.n_new[y]; f += per_block) {
            ev_block.push_back((int32_t)y);
            ev_block.push_back((int32_t)f);
        }
    }
    year_blk[Y] = (int32_t)(ev_block.size() / 2);
    const int64_t nE = (int64_t)ev_block.size() / 2;
    const int64_t nP = lik_batch_paper_blocks(q.n_papers, B);
    const int64_t part = ((Y + 2) * nP + 4 * nE) * B + 1;
    // segment-sum jobs from the per-paper offsets of the sorted slots: runs
    // of whole papers up to `small` positions, larger papers in pieces
    std::vector<int32_t> so(q.n_papers + 1);
    if (q.n_papers >= 0)
        PG_HIP(hipMemcpy(so.data(), q.seg_off, (q.n_papers + 1) * 4, hipMemcpyDeviceToHost));
    const int64_t small = lik_batch_small_job(B);
    std::vector<int64_t> job_start;
    std::vector<int32_t> job_tag, big_paper, big_first;
    int32_t pieces = 0;
    int64_t cur = 0, cur_len = 0;
    for (int64_t j = 0; j < q.n_papers; ++j) {
        const int64_t len = (int64_t)so[j + 1] - so[j];
        if (len == 0) continue;
        if (len > small) {
            if (cur_len > 0) { job_start.push_back(cur); job_tag.push_back(-1); }
            big_paper.push_back((int32_t)j);
            big_first.push_back(pieces);
            for (int64_t o = 0; o < len; o += kLikBBigPiece) {
                job_start.push_back(so[j] + o);
                job_tag.push_back(pieces++);
            }
            cur = so[j + 1];
            cur_len = 0;
        } else {
            if (cur_len + len > small) {
                job_start.push_back(cur);
                job_tag.push_back(-1);
                cur = so[j];
                cur_len = 0;
            }
            cur_len += len;
        }
    }
    if (cur_len > 0) { job_start.push_back(cur); job_tag.push_back(-1); }
    const int64_t n_jobs = (int64_t)job_tag.size();
    job_start.push_back(q.filled);
    big_first.push_back(pieces);
    hipError_t r = dalloc(&q.b_ev_block, std::max<int64_t>(2 * nE, 1));
    if (r == hipSuccess) r = dalloc(&q.b_year_blk, Y + 1);
    if (r == hipSuccess) r = dalloc(&q.b_part, part);
    if (r == hipSuccess) r = dalloc(&q.b_year_sums, std::max<int64_t>(Y, 1) * kLikYearCols * B);
    if (r == hipSuccess) r = dalloc(&q.b_share, std::max<int64_t>(q.filled, 1) * B);
    if (r == hipSuccess) r = dalloc(&q.b_evt, std::max<int64_t>(q.n_papers, 1) * B);
    if (r == hipSuccess) r = dalloc(&q.b_job_start, n_jobs + 1);
    if (r == hipSuccess) r = dalloc(&q.b_job_tag, std::max<int64_t>(n_jobs, 1));
    if (r == hipSuccess) r = dalloc(&q.b_bigpart, (int64_t)std::max<int32_t>(pieces, 1) * B);
    if (r == hipSuccess) r = dalloc(&q.b_big_paper, std::max<int64_t>((int64_t)big_paper.size(), 1));
    if (r == hipSuccess) r = dalloc(&q.b_big_first, (int64_t)big_first.size());
    if (r == hipSuccess)
        r = hipMemcpy(q.b_job_start, job_start.data(), (n_jobs + 1) * 8, hipMemcpyHostToDevice);
    if (r == hipSuccess && n_jobs > 0)
        r = hipMemcpy(q.b_job_tag, job_tag.data(), n_jobs * 4, hipMemcpyHostToDevice);
    if (r == hipSuccess && !big_paper.empty())
        r = hipMemcpy(q.b_big_paper, big_paper.data(), big_paper.size() * 4, hipMemcpyHostToDevice);
    if (r == hipSuccess)
        r = hipMemcpy(q.b_big_first, big_first.data(), big_first.size() * 4, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = dalloc(&q.b_at, 2 * B);
    if (r == hipSuccess) r = dalloc(&q.b_out, 3 * B);
    if (r == hipSuccess && nE > 0)
        r = hipMemcpy(q.b_ev_block, ev_block.data(), nE * 8, hipMemcpyHostToDevice);
    if (r == hipSuccess)
        r = hipMemcpy(q.b_year_blk, year_blk.data(), (Y + 1) * 4, hipMemcpyHostToDevice);
    if (r != hipSuccess) {
        seq_free_batch(q);
        return fail(ctx, r == hipErrorOutOfMemory ? PG_ENOMEM : PG_EHIP,
                    std::string("pg_pa_loglik_grad_batch: buffers for ") + std::to_string(B) +
                        " points: " + hipGetErrorString(r));
    }
    q.bB = B;
    std::copy(key, key + 4, q.b_key);
    q.b_shape = sh;
    q.b_nE = nE;
    q.b_nP = nP;
    q.b_n_jobs = n_jobs;
    q.b_n_big = (int32_t)big_paper.size();
    return PG_OK;
}

// Algorithmic HBM bytes of one batched evaluation at B points (the formula
// in paragraph.h).
double lik_batch_bytes(const pg_seq_store& q, int B) {
    const double PY = (double)q.paper_years, P = (double)q.n_papers, NP = (double)q.new_total,
                 S = (double)q.filled;
    return 8.0 * PY + 12.0 * P + 4.0 * NP + 21.0 * S + (double)B * (32.0 * P + 20.0 * S);
}

int lik_eval(pg_ctx* ctx, const pg_lik_params* p, pg_lik_out* out, bool grad) {
    const char* fn = grad ? "pg_pa_loglik_grad" : "pg_pa_loglik";
    if (!ctx) return PG_EINVAL;
    if (!p || p->size != sizeof(pg_lik_params) || !out || out->size != sizeof(pg_lik_out))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": null argument or struct size mismatch");
    pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, std::string(fn) + ": no sequence");
    if (!p->alpha_tau_dev && !(std::isfinite(p->alpha) && p->tau > 0.0 && std::isfinite(p->tau)))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": alpha must be finite and tau > 0");
    if (q.n_papers > 0 && !p->logfit_dev)
        return fail(ctx, PG_EINVAL, std::string(fn) + ": logfit_dev is NULL");
    if (grad && q.n_papers > 0 && !out->grad_logfit_dev)
        return fail(ctx, PG_EINVAL, std::string(fn) + ": grad_logfit_dev is NULL");
    for (size_t y = 0; y < q.n_exist.size(); ++y)
        if (q.n_exist[y] + q.n_new[y] > q.n_papers)
            return fail(ctx, PG_ESTATE, std::string(fn) + ": births not set for every paper");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    {
        int rc = seq_prepare(ctx);
        if (rc) return rc;
    }
    hipStream_t s = p->stream_given ? static_cast<hipStream_t>(p->stream) : ctx->stream;
    const int grp = grad ? PG_K_LIK_GRAD : PG_K_LIK;
    const bool prof = ctx->prof && ((ctx->prof_mask >> grp) & 1u);
    if (prof) PG_HIP(hipEventRecord(ctx->pev[grp][0], s));
    const LikSeq v = seq_view(q);
    LikArgs a;
    a.alpha = p->alpha;
    a.tau = p->tau;
    a.at_dev = p->alpha_tau_dev;
    a.logfit = p->logfit_dev;
    const int64_t nP = q.n_paper_blocks, nE = q.n_ev_blocks, Y = v.n_years;
    double* part_p = q.part;
    double* part_e = q.part + Y * nP;
    double* part_b = part_e + 4 * nE;
    launch_lik_forward_sums(s, v, a, nP, part_p, q.year_sums);
    launch_lik_slots(s, grad, q.cite_shape, nE, q.d_ev_block, q.d_ev_year_blk, v, a, q.year_sums,
                     q.share, part_e);
    if (grad) {
        const LikSegs sg{q.seg_off, q.evt, q.carry};
        launch_lik_backward(s, v, a, nP, q.year_sums, q.filled, q.skey, q.share, sg,
                            out->grad_logfit_dev, part_b);
    }
    launch_lik_final(s, v.n_years, grad, nP, part_b, a, q.year_sums, q.d_scalars,
                     out->scalars_dev);
    PG_HIP(hipGetLastError());
    if (prof) {
        PG_HIP(hipEventRecord(ctx->pev[grp][1], s));
        PG_HIP(hipEventSynchronize(ctx->pev[grp][1]));
        float ms = 0.0f;
        PG_HIP(hipEventElapsedTime(&ms, ctx->pev[grp][0], ctx->pev[grp][1]));
        ctx->prof_ms[grp] += ms;
        ctx->prof_bytes[grp] += lik_bytes(q, grad);
        ctx->prof_calls[grp] += 1;
    }
    if (out->host3) {
        PG_HIP(hipMemcpyAsync(out->host3, q.d_scalars, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        PG_HIP(hipStreamSynchronize(s));
    }
    return PG_OK;
}

}  // namespace

void seq_release(pg_ctx* ctx) {
    pg_seq_store& q = ctx->seq;
    free_ptr(q.deg); free_ptr(q.cit); free_ptr(q.kind); free_ptr(q.cnt); free_ptr(q.birth);
    q.deg = q.cit = q.cnt = q.birth = nullptr;
    q.kind = nullptr;
    seq_free_tables(q);
    seq_clear(q);
    q.reserved = false;
}

int seq_record_snapshot(pg_ctx* ctx, int32_t year, int64_t n_exist, int64_t n_new, int32_t m,
                        int model) {
    ctx->seq.pending = false;
    if (!ctx->seq_recording) return PG_OK;
    if (model != PG_MODEL_PA)
        return fail(ctx, PG_ESTATE, "pg_year_step: sequence recording holds PA years only");
    if (m != ctx->seq.m)
        return fail(ctx, PG_EINVAL, "pg_year_step: m differs from the recorded sequence's m");
    {
        int rc = seq_check_year(ctx, "pg_year_step (recording)", n_exist, n_new);
        if (rc) return rc;
    }
    pg_seq_store& q = ctx->seq;
    if (n_exist > 0)
        PG_HIP(hipMemcpyAsync(q.deg + q.paper_years, ctx->in_deg, n_exist * 4,
                              hipMemcpyDeviceToDevice, ctx->stream));
    q.pending_year = year;
    q.pending_n_exist = n_exist;
    q.pending_n_new = n_new;
    q.pending = true;
    return PG_OK;
}

int seq_record_draws(pg_ctx* ctx, int64_t filled) {
    pg_seq_store& q = ctx->seq;
    if (!ctx->seq_recording || !q.pending) return PG_OK;
    q.pending = false;
    const int64_t D = q.pending_n_new, N = q.pending_n_exist, m = q.m;
    hipStream_t s = ctx->stream;
    if (D > 0 && m > 0) {
        PG_HIP(hipMemcpyAsync(q.cit + q.new_total * m, ctx->cit, D * m * 4,
                              hipMemcpyDeviceToDevice, s));
        PG_HIP(hipMemcpyAsync(q.kind + q.new_total * m, ctx->kind, D * m,
                              hipMemcpyDeviceToDevice, s));
    }
    if (D > 0)
        PG_HIP(hipMemcpyAsync(q.cnt + q.new_total, ctx->cnt, D * 4, hipMemcpyDeviceToDevice, s));
    if (N + D > 0)
        PG_HIP(hipMemcpyAsync(q.birth, ctx->birth, (N + D) * 4, hipMemcpyDeviceToDevice, s));
    seq_commit_year(q, q.pending_year, N, D);
    q.filled += filled;
    q.n_papers = std::max<int64_t>(q.n_papers, N + D);
    return PG_OK;
}

extern "C" {

int pg_seq_reserve(pg_ctx* ctx, int32_t m, int64_t max_paper_years, int64_t max_new_papers,
                   int64_t max_papers) {
    if (!ctx) return PG_EINVAL;
    if (m < 0 || m > kMaxCites || max_paper_years < 0 || max_new_papers < 0 || max_papers < 0)
        return fail(ctx, PG_EINVAL, "pg_seq_reserve: need 0 <= m <= 64 and non-negative sizes");
    if (max_new_papers * std::max<int64_t>(m, 1) >= (int64_t(1) << 31) ||
        max_papers >= (int64_t(1) << 31) - 1)
        return fail(ctx, PG_ECAP, "pg_seq_reserve: slots and papers must stay below 2^31");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    PG_HIP(hipStreamSynchronize(ctx->stream));
    seq_release(ctx);
    pg_seq_store& q = ctx->seq;
    hipError_t r = dalloc(&q.deg, max_paper_years);
    if (r == hipSuccess) r = dalloc(&q.cit, max_new_papers * std::max<int64_t>(m, 1));
    if (r == hipSuccess) r = dalloc(&q.kind, max_new_papers * std::max<int64_t>(m, 1));
    if (r == hipSuccess) r = dalloc(&q.cnt, max_new_papers);
    if (r == hipSuccess) r = dalloc(&q.birth, max_papers);
    if (r != hipSuccess) {
        seq_release(ctx);
        return fail(ctx, PG_ENOMEM, "pg_seq_reserve: device allocation failed");
    }
    q.m = m;
    q.cap_paper_years = max_paper_years;
    q.cap_new = max_new_papers;
    q.cap_papers = max_papers;
    q.reserved = true;
    return PG_OK;
}

int pg_seq_record(pg_ctx* ctx, int on) {
    if (!ctx) return PG_EINVAL;
    if (on && !ctx->seq.reserved) return fail(ctx, PG_ESTATE, "pg_seq_record: no sequence reserved");
    ctx->seq_recording = on != 0;
    return PG_OK;
}

int pg_seq_append_year(pg_ctx* ctx, int32_t year, int64_t n_exist, int64_t n_new,
                       const int32_t* deg, const int32_t* cit, const int8_t* kind,
                       const int32_t* cnt) {
    if (!ctx) return PG_EINVAL;
    {
        int rc = seq_check_year(ctx, "pg_seq_append_year", n_exist, n_new);
        if (rc) return rc;
    }
    pg_seq_store& q = ctx->seq;
    const int64_t m = q.m;
    if ((n_exist > 0 && !deg) || (n_new > 0 && (!cnt || (m > 0 && (!cit || !kind)))))
        return fail(ctx, PG_EINVAL, "pg_seq_append_year: null array");
    for (int64_t j = 0; j < n_exist; ++j)
        if (deg[j] < 0) return fail(ctx, PG_EINVAL, "pg_seq_append_year: negative in-degree");
    for (int64_t i = 0; i < n_new; ++i) {
        if (cnt[i] < 0 || cnt[i] > m)
            return fail(ctx, PG_EINVAL, "pg_seq_append_year: count outside [0, m]");
        for (int64_t sl = 0; sl < cnt[i]; ++sl) {
            const int32_t t = cit[i * m + sl];
            const int8_t k = kind[i * m + sl];
            if (t < 0 || t >= n_exist)
                return fail(ctx, PG_EINVAL, "pg_seq_append_year: target outside [0, n_exist)");
            if (k < 0 || k > 2)
                return fail(ctx, PG_EINVAL, "pg_seq_append_year: kind outside {0,1,2}");
        }
    }
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (n_exist > 0)
        PG_HIP(hipMemcpy(q.deg + q.paper_years, deg, n_exist * 4, hipMemcpyHostToDevice));
    if (n_new > 0 && m > 0) {
        PG_HIP(hipMemcpy(q.cit + q.new_total * m, cit, n_new * m * 4, hipMemcpyHostToDevice));
        PG_HIP(hipMemcpy(q.kind + q.new_total * m, kind, n_new * m, hipMemcpyHostToDevice));
    }
    if (n_new > 0) PG_HIP(hipMemcpy(q.cnt + q.new_total, cnt, n_new * 4, hipMemcpyHostToDevice));
    for (int64_t i = 0; i < n_new; ++i) q.filled += cnt[i];
    seq_commit_year(q, year, n_exist, n_new);
    return PG_OK;
}

int pg_seq_set_birth(pg_ctx* ctx, int64_t n_papers, const int32_t* birth) {
    if (!ctx) return PG_EINVAL;
    pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, "pg_seq_set_birth: no sequence reserved");
    if (n_papers < 0 || n_papers > q.cap_papers || (n_papers > 0 && !birth))
        return fail(ctx, PG_EINVAL, "pg_seq_set_birth: bad size or null array");
    for (size_t y = 0; y < q.n_exist.size(); ++y)
        if (q.n_exist[y] + q.n_new[y] > n_papers)
            return fail(ctx, PG_EINVAL, "pg_seq_set_birth: fewer papers than the sequence uses");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (n_papers > 0) PG_HIP(hipMemcpy(q.birth, birth, n_papers * 4, hipMemcpyHostToDevice));
    q.n_papers = n_papers;
    q.tables_dirty = true;
    return PG_OK;
}

int pg_seq_get_info(pg_ctx* ctx, pg_seq_info* info) {
    if (!ctx) return PG_EINVAL;
    if (!info || info->size != sizeof(pg_seq_info))
        return fail(ctx, PG_EINVAL, "pg_seq_get_info: null info or struct size mismatch");
    const pg_seq_store& q = ctx->seq;
    info->n_years = (int32_t)q.n_exist.size();
    info->m = q.m;
    info->n_papers = q.n_papers;
    info->paper_years = q.paper_years;
    info->new_papers = q.new_total;
    return PG_OK;
}

int pg_seq_download(pg_ctx* ctx, int64_t* n_exist, int64_t* n_new, int32_t* year_of,
                    int32_t* deg, int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* birth) {
    if (!ctx) return PG_EINVAL;
    const pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, "pg_seq_download: no sequence reserved");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    PG_HIP(hipStreamSynchronize(ctx->stream));
    const size_t Y = q.n_exist.size();
    if (n_exist && Y) std::memcpy(n_exist, q.n_exist.data(), Y * 8);
    if (n_new && Y) std::memcpy(n_new, q.n_new.data(), Y * 8);
    if (year_of && Y) std::memcpy(year_of, q.year_of.data(), Y * 4);
    const int64_t m = q.m;
    if (deg && q.paper_years)
        PG_HIP(hipMemcpy(deg, q.deg, q.paper_years * 4, hipMemcpyDeviceToHost));
    if (cit && q.new_total * m)
        PG_HIP(hipMemcpy(cit, q.cit, q.new_total * m * 4, hipMemcpyDeviceToHost));
    if (kind && q.new_total * m)
        PG_HIP(hipMemcpy(kind, q.kind, q.new_total * m, hipMemcpyDeviceToHost));
    if (cnt && q.new_total) PG_HIP(hipMemcpy(cnt, q.cnt, q.new_total * 4, hipMemcpyDeviceToHost));
    if (birth && q.n_papers)
        PG_HIP(hipMemcpy(birth, q.birth, q.n_papers * 4, hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_pa_loglik_grad_batch(pg_ctx* ctx, const pg_lik_batch_params* p, pg_lik_batch_out* out) {
    const char* fn = "pg_pa_loglik_grad_batch";
    if (!ctx) return PG_EINVAL;
    if (!p || p->size != sizeof(pg_lik_batch_params) || !out ||
        out->size != sizeof(pg_lik_batch_out))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": null argument or struct size mismatch");
    pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, std::string(fn) + ": no sequence");
    const int B = p->n_points;
    if (!lik_batch_ok(B))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": n_points must be 4, 8 or 16");
    if (!lik_batch_shape_ok(q.m, B))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": " + std::to_string(q.m) +
                                        " slots per paper do not fit the slot pass at " +
                                        std::to_string(B) + " points (use fewer points)");
    if ((q.n_papers + 1) * (int64_t)B >= (int64_t(1) << 31))
        return fail(ctx, PG_ECAP, std::string(fn) + ": papers times points must stay below 2^31");
    if (p->arith != PG_LIK_ARITH_SIM && p->arith != PG_LIK_ARITH_FAST)
        return fail(ctx, PG_EINVAL, std::string(fn) + ": arith must be PG_LIK_ARITH_SIM or _FAST");
    if (!p->alpha_tau_dev) {
        if (!p->alpha_tau) return fail(ctx, PG_EINVAL, std::string(fn) + ": no alpha/tau given");
        for (int b = 0; b < B; ++b)
            if (!(std::isfinite(p->alpha_tau[2 * b]) && p->alpha_tau[2 * b + 1] > 0.0 &&
                  std::isfinite(p->alpha_tau[2 * b + 1])))
                return fail(ctx, PG_EINVAL, std::string(fn) + ": alpha must be finite and tau > 0");
    }
    if (q.n_papers > 0 && (!p->logfit_dev || !out->grad_logfit_dev))
        return fail(ctx, PG_EINVAL, std::string(fn) + ": logfit_dev or grad_logfit_dev is NULL");
    for (size_t y = 0; y < q.n_exist.size(); ++y)
        if (q.n_exist[y] + q.n_new[y] > q.n_papers)
            return fail(ctx, PG_ESTATE, std::string(fn) + ": births not set for every paper");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    if (int rc = seq_prepare(ctx)) return rc;
    if (int rc = batch_prepare(ctx, B)) return rc;
    hipStream_t s = p->stream_given ? static_cast<hipStream_t>(p->stream) : ctx->stream;
    const double* at = p->alpha_tau_dev;
    if (!at) {
        PG_HIP(hipMemcpyAsync(q.b_at, p->alpha_tau, 2 * B * sizeof(double), hipMemcpyHostToDevice,
                              s));
        at = q.b_at;
    }
    const bool prof = ctx->prof && ((ctx->prof_mask >> PG_K_LIK_BATCH) & 1u);
    if (prof) PG_HIP(hipEventRecord(ctx->pev[PG_K_LIK_BATCH][0], s));
    LikBatchBufs bb;
    bb.at = at;
    bb.logfit = p->logfit_dev;
    bb.year_sums = q.b_year_sums;
    bb.part = q.b_part;
    bb.ev_block = q.b_ev_block;
    bb.year_blk = q.b_year_blk;
    bb.nblk_p = q.b_nP;
    bb.nblk_e = q.b_nE;
    bb.share = q.b_share;
    bb.S = q.filled;
    bb.skey = q.skey;
    bb.seg_off = q.seg_off;
    bb.evt = q.b_evt;
    bb.job_start = q.b_job_start;
    bb.job_tag = q.b_job_tag;
    bb.n_jobs = q.b_n_jobs;
    bb.bigpart = q.b_bigpart;
    bb.big_paper = q.b_big_paper;
    bb.big_first = q.b_big_first;
    bb.n_big = q.b_n_big;
    bb.grad = out->grad_logfit_dev;
    bb.out = q.b_out;
    bb.out2 = out->scalars_dev;
    launch_lik_batch(s, B, seq_view(q), bb, q.b_shape, p->arith == PG_LIK_ARITH_FAST);
    PG_HIP(hipGetLastError());
    if (prof) {
        PG_HIP(hipEventRecord(ctx->pev[PG_K_LIK_BATCH][1], s));
        PG_HIP(hipEventSynchronize(ctx->pev[PG_K_LIK_BATCH][1]));
        float ms = 0.0f;
        PG_HIP(hipEventElapsedTime(&ms, ctx->pev[PG_K_LIK_BATCH][0], ctx->pev[PG_K_LIK_BATCH][1]));
        ctx->prof_ms[PG_K_LIK_BATCH] += ms;
        ctx->prof_bytes[PG_K_LIK_BATCH] += lik_batch_bytes(q, B);
        ctx->prof_calls[PG_K_LIK_BATCH] += 1;
    }
    if (out->host) {
        PG_HIP(hipMemcpyAsync(out->host, q.b_out, 3 * B * sizeof(double), hipMemcpyDeviceToHost, s));
        PG_HIP(hipStreamSynchronize(s));
    }
    return PG_OK;
}

int pg_pa_loglik(pg_ctx* ctx, const pg_lik_params* p, pg_lik_out* out) {
    return lik_eval(ctx, p, out, false);
}

int pg_pa_loglik_grad(pg_ctx* ctx, const pg_lik_params* p, pg_lik_out* out) {
    return lik_eval(ctx, p, out, true);
}

}  // extern "C"

extern "C" {

int pg_lik_autotune(pg_ctx* ctx, const pg_lik_params* p, int32_t reps, pg_tune_report* report) {
    if (!ctx) return PG_EINVAL;
    if (!p || p->size != sizeof(pg_lik_params) || (report && report->size != sizeof(pg_tune_report)))
        return fail(ctx, PG_EINVAL, "pg_lik_autotune: null argument or struct size mismatch");
    if (reps < 1 || reps > 64) return fail(ctx, PG_EINVAL, "pg_lik_autotune: reps in [1, 64]");
    pg_seq_store& q = ctx->seq;
    if (!q.reserved) return fail(ctx, PG_ESTATE, "pg_lik_autotune: no sequence");
    PG_HIP(hipSetDevice(ctx->cfg.device));
    pg_lik_params pp = *p;
    pp.stream = nullptr;
    pp.stream_given = 0;  // the context's stream, timed by its events
    double* grad = nullptr;
    if (dalloc(&grad, std::max<int64_t>(q.n_papers, 1)) != hipSuccess)
        return fail(ctx, PG_ENOMEM, "pg_lik_autotune: device allocation failed");
    struct Freer { double* g; ~Freer() { (void)hipFree(g); } } freer{grad};
    const int saved = ctx->tune_lik_lanes;
    const int def_lanes = lik_cite_shape(q.m, 0).G;
    std::vector<int> cand;
    std::vector<double> ms;
    double ref[3] = {0, 0, 0}, max_dev = 0.0, default_ms = 0.0;
    for (int lanes : {1, 2, 4, 8}) {
        if (!lik_lanes_ok(q.m, lanes)) continue;
        if (lanes != ctx->tune_lik_lanes) q.tables_dirty = true;
        ctx->tune_lik_lanes = lanes;
        std::vector<double> v;
        for (int r = 0; r <= reps; ++r) {
            double h[3];
            pg_lik_out out{};
            out.size = sizeof out;
            out.grad_logfit_dev = grad;
            out.host3 = r == 0 ? h : nullptr;
            PG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            if (int rc = pg_pa_loglik_grad(ctx, &pp, &out)) {
                ctx->tune_lik_lanes = saved;
                q.tables_dirty = true;
                return rc;
            }
            PG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
            PG_HIP(hipEventSynchronize(ctx->ev1));
            if (r == 0) {  // untimed: table rebuild; results compared
                if (cand.empty()) {
                    for (int k = 0; k < 3; ++k) ref[k] = h[k];
                } else {
                    for (int k = 0; k < 3; ++k) {
                        const double d = std::fabs(h[k] - ref[k]) / std::max(std::fabs(ref[k]), 1e-300);
                        max_dev = std::max(max_dev, d);
                    }
                }
                continue;
            }
            float t = 0.0f;
            PG_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
            v.push_back(t);
        }
        std::sort(v.begin(), v.end());
        cand.push_back(lanes);
        ms.push_back(v.size() % 2 ? v[v.size() / 2] : 0.5 * (v[v.size() / 2 - 1] + v[v.size() / 2]));
        if (lanes == def_lanes) default_ms = ms.back();
    }
    const size_t best = std::min_element(ms.begin(), ms.end()) - ms.begin();
    if (ctx->tune_lik_lanes != cand[best]) q.tables_dirty = true;
    ctx->tune_lik_lanes = cand[best];
    if (report) {
        report->n_trials = (int32_t)cand.size();
        for (size_t k = 0; k < cand.size(); ++k) {
            pg_tuning t{};
            t.size = sizeof t;
            t.pa_lanes = ctx->tune_pa_lanes;
            t.lik_lanes = cand[k];
            t.bfs_mode = ctx->tune_bfs_mode;
            t.bfs_small_max = ctx->tune_small_max;
            t.bfs_mid_max = ctx->tune_mid_max;
            t.merge_out = ctx->tune_merge_out;
            report->trial[k] = t;
            report->trial_ms[k] = ms[k];
        }
        report->chosen = report->trial[best];
        report->chosen_ms = ms[best];
        report->default_ms = default_ms;
        report->max_rel_dev = max_dev;
    }
    return PG_OK;
}

}  // extern "C"
