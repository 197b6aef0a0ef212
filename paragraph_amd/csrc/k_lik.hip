// k_lik.hip -- PA log-likelihood of a recorded citation sequence and its
// gradient with respect to alpha, tau and every paper's log-fitness.  No
// reference counterpart: this is the differentiable view of the year step's
// PA draws (k_pa.hip), with weights frozen per year.
//
//   x_j(y) = alpha*log(k_j(y)+1) - (year_y - b_j)/tau + phi_j   (FP32, the
//   w_j(y) = exp(x_j(y))                                          simulator's
//                                                                 arithmetic)
//   Z_y    = sum_{j < N_y} w_j(y)                                 (FP64)
//   log L  = sum over PA slots (kinds 0, 2) of x_t - log(Z_y - P)  (FP64)
// where P sums the weights of the slots the same paper filled before (all
// kinds: a neighbourhood pick also leaves the PA partition).
//
// Gradient (A_y = sum of 1/D over year y's PA slots, D = Z_y - P):
//   dL/dalpha = sum_slots [L_t + PL/D] - sum_y A_y * sum_j w_j L_j
//   dL/dtau   = sum_slots [age_t + Pa/D]/tau^2 - sum_y A_y * sum_j w_j age_j / tau^2
//   dL/dphi_j = #PA slots citing j + sum_slots citing j of w_j * later
//               - sum_y A_y w_j(y)
// with PL, Pa the w*L and w*age sums over earlier slots and `later` the sum
// of 1/D over the same paper's later PA slots.
//
// Passes of one evaluation.  Every reduction has a fixed order, so log L
// and the whole gradient are bit-reproducible from call to call:
//   k_lik_papers<false>  paper-major, kLikPPT papers per thread: per block
//                        and year, the partial Z_y
//   k_lik_year_z         Z_y = fixed-order sum of the block partials
//   k_lik_cite<GRAD,G,K> G lanes per citing paper, K slots per lane:
//                        D = Z_y - P per PA slot; per block partials of
//                        log L (and A, alpha, tau terms); GRAD: each slot's
//                        dL/dphi_t share [PA slot] + w_t * later, stored at
//                        the slot's position in the by-target order
//   k_lik_year_ev        per year: log L, A_y and the slot terms
//   k_lik_segsum         the shares summed per target paper over the slots
//                        sorted by target (sorted once per sequence): runs
//                        inside a 2048-slot chunk are final, runs crossing a
//                        chunk edge leave per-chunk carries
//   k_lik_papers<true>   recomputes w_j(y) (never stored): dL/dphi_j =
//                        slot shares - sum_y A_y w_j(y); partials of the
//                        A-weighted sums
//   k_lik_final          combines into {log L, dL/dalpha, dL/dtau}
#include <hip/hip_runtime.h>

#include "pg_internal.h"
#include "pg_math.h"

namespace pg {

namespace {

constexpr int kLikBT = 256;             // threads per block, every pass
constexpr int kLikWaves = kLikBT / 64;

struct LikScal {
    float alpha_f, inv_tau_f;
    double tau, itau2;
};

// alpha/tau from the host values, or from a device pair {alpha, tau}.
__device__ inline LikScal lik_scal(const LikArgs& a) {
    LikScal s;
    double alpha = a.alpha, tau = a.tau;
    if (a.at_dev) {
        alpha = a.at_dev[0];
        tau = a.at_dev[1];
    }
    s.alpha_f = (float)alpha;
    s.inv_tau_f = (float)(1.0 / tau);
    s.tau = tau;
    s.itau2 = 1.0 / (tau * tau);
    return s;
}

// The simulator's weight arithmetic (pa_weight) from L = log(k+1), which is
// precomputed per paper-year and per slot with the simulator's det_logf, and
// the age (as float, exact below 2^24); keeps x = log w.
__device__ inline float lik_weight_la(float L, float age, float phi, const LikScal& s, float& x) {
    const float t1 = __builtin_fmaf(-age, s.inv_tau_f, phi);
    x = __builtin_fmaf(s.alpha_f, L, t1);
    return det_expf(x);
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Fixed-order block sum (wave butterflies, then waves in order); result in
// every thread.
__device__ inline double block_sum_d(double v, double* red) {
    v = wave_sum_d(v);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kLikWaves; ++w) t += red[w];
    return t;
}

// Slot shares of paper j from the segment sums (k_lik_segsum's outputs).
__device__ inline double slot_share(int64_t j, const int32_t* __restrict__ seg_off,
                                    const double* __restrict__ evt,
                                    const double* __restrict__ carry) {
    const int64_t o0 = seg_off[j], o1 = seg_off[j + 1];
    if (o0 == o1) return 0.0;
    const int64_t b0 = o0 / kLikSegChunk, b1 = (o1 - 1) / kLikSegChunk;
    if (b0 == b1) return evt[j];
    double e = carry[2 * b0 + 1];
    for (int64_t b = b0 + 1; b < b1; ++b) e += carry[2 * b];
    return e + carry[2 * b1];
}

// Paper-major pass.  Thread t of block B holds papers
// j = B*kLikBT*kLikPPT + r*kLikBT + t (r < kLikPPT) and loops over the years
// in which they exist (j < N_y).
//   BWD = false: part[y * nblk + B] = sum of the block's w_j(y).
//   BWD = true : grad_phi[j] = slot shares - sum_y A_y w_j(y);
//                part[B], part[nblk + B] = sums of A_y w L and A_y w age.
template <bool BWD>
__global__ __launch_bounds__(kLikBT) void k_lik_papers(LikSeq q, LikArgs a,
                                                       const double* __restrict__ year_sums,
                                                       LikSegs sg, double* __restrict__ grad_phi,
                                                       double* __restrict__ part) {
    __shared__ double s_red[kLikWaves][kLikMaxYears];
    __shared__ double s_red2[kLikWaves];
    const LikScal s = lik_scal(a);
    const int64_t nblk = gridDim.x;
    const int64_t j0 = (int64_t)blockIdx.x * (kLikBT * kLikPPT);
    const int wave = threadIdx.x >> 6;
    float phi[kLikPPT];
    int32_t b[kLikPPT];
    double g[kLikPPT];
#pragma unroll
    for (int r = 0; r < kLikPPT; ++r) {
        const int64_t j = j0 + r * kLikBT + threadIdx.x;
        phi[r] = 0.0f;
        b[r] = 0;
        g[r] = 0.0;
        if (j < q.n_papers) {
            phi[r] = a.logfit[j];
            b[r] = q.birth[j];
        }
    }
    double gL = 0.0, ga = 0.0;
    // kLikYU years at a time: their log(k+1) loads are issued together
    // before any is used (the sums still run year by year, in order)
    constexpr int kLikYU = 4;
    for (int y0 = 0; y0 < q.n_years; y0 += kLikYU) {
    float Lv[kLikYU][kLikPPT];
#pragma unroll
    for (int yy = 0; yy < kLikYU; ++yy) {
        const int y = y0 + yy;
        const int64_t Ny = y < q.n_years ? q.n_exist[y] : 0;
        const int64_t doff = y < q.n_years ? q.deg_off[y] : 0;
#pragma unroll
        for (int r = 0; r < kLikPPT; ++r) {
            const int64_t j = j0 + r * kLikBT + threadIdx.x;
            Lv[yy][r] = j < Ny ? q.Lpy[doff + j] : 0.0f;
        }
    }
#pragma unroll
    for (int yy = 0; yy < kLikYU; ++yy) {
        const int y = y0 + yy;
        if (y >= q.n_years) break;
        const int64_t Ny = q.n_exist[y];
        if (j0 >= Ny) {                  // block-uniform: no paper of the block exists
            if (!BWD && threadIdx.x == 0) part[(int64_t)y * nblk + blockIdx.x] = 0.0;
            continue;
        }
        const int32_t year = q.year_of[y];
        const double Ay = BWD ? year_sums[y * kLikYearCols + 1] : 0.0;
        double zacc = 0.0;
#pragma unroll
        for (int r = 0; r < kLikPPT; ++r) {
            const int64_t j = j0 + r * kLikBT + threadIdx.x;
            if (j < Ny) {
                const float L = Lv[yy][r];
                const int32_t age = year - b[r];
                float x;
                const double wv = (double)lik_weight_la(L, (float)age, phi[r], s, x);
                if (BWD) {
                    const double aw = Ay * wv;
                    g[r] += aw;
                    gL += aw * (double)L;
                    ga += aw * (double)age;
                } else {
                    zacc += wv;
                }
            }
        }
        if (!BWD) {
            const double ws = wave_sum_d(zacc);
            if ((threadIdx.x & 63) == 0) s_red[wave][y] = ws;
        }
    }
    }
    if (BWD) {
#pragma unroll
        for (int r = 0; r < kLikPPT; ++r) {
            const int64_t j = j0 + r * kLikBT + threadIdx.x;
            if (j < q.n_papers) grad_phi[j] = slot_share(j, sg.seg_off, sg.evt, sg.carry) - g[r];
        }
        const double tL = block_sum_d(gL, s_red2);
        const double tA = block_sum_d(ga, s_red2);
        if (threadIdx.x == 0) {
            part[blockIdx.x] = tL;
            part[nblk + blockIdx.x] = tA;
        }
    } else {
        __syncthreads();
        for (int y = threadIdx.x; y < q.n_years; y += kLikBT) {
            if (j0 >= q.n_exist[y]) continue;
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kLikWaves; ++w) t += s_red[w][y];
            part[(int64_t)y * nblk + blockIdx.x] = t;
        }
    }
}

// Z_y = sum over the paper blocks of part[y * nblk + b], fixed order.
__global__ __launch_bounds__(kLikBT) void k_lik_year_z(int64_t nblk, const double* __restrict__ part,
                                                        double* __restrict__ year_sums) {
    __shared__ double red[kLikWaves];
    const int y = blockIdx.x;
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += kLikBT) v += part[(int64_t)y * nblk + b];
    const double t = block_sum_d(v, red);
    if (threadIdx.x == 0) year_sums[y * kLikYearCols + 0] = t;
}

// Slot pass.  G lanes per citing paper (G consecutive lanes of a wave),
// lane r owning the contiguous slots [r*K, r*K + K) of its paper; block B
// holds kLikBT/G consecutive citing papers of one year starting at
// ev_block[2B+1] (year ev_block[2B]).
//   loop 1  per slot, independent: target, gathers, FP32 weight, x, L, age
//           (a slot past the count re-reads the paper's slot 0, which is
//           valid, and is masked off: no branches, the loads overlap);
//   chain   the draw-order prefix P of the weights, added slot by slot in
//           the oracle's order: lane r starts from lane r-1's end value;
//   loop 2  D = Z_y - P per PA slot: log L, and (GRAD) 1/D;
//   GRAD    later(q) = sum of 1/D over the paper's PA slots after q (suffix
//           over the lanes, then within the lane in reverse); each slot's
//           dL/dphi share [PA slot] + w_t * later goes to its by-target
//           position share[pos_t[slot]]; the alpha/tau prefix terms are
//           sum_j PL_j/D_j = sum_i w_i L_i later_i (likewise for age).
// Partials per block, column c at part[c * nblk + B]: 0 log L, 1 A,
// 2 alpha terms, 3 tau terms.
template <bool GRAD, int G, int K>
__global__ __launch_bounds__(kLikBT) void k_lik_cite(LikSeq q, LikArgs a,
                                                     const int32_t* __restrict__ ev_block,
                                                     const double* __restrict__ year_sums,
                                                     double* __restrict__ share,
                                                     double* __restrict__ part) {
    __shared__ double red[kLikWaves];
    const LikScal s = lik_scal(a);
    const int64_t nblk = gridDim.x;
    const int y = ev_block[2 * blockIdx.x];
    const int lg = threadIdx.x % G;                   // lane within the paper's group
    const int64_t i = (int64_t)ev_block[2 * blockIdx.x + 1] + threadIdx.x / G;
    const int64_t Dy = q.n_new[y];
    const double Z = year_sums[y * kLikYearCols + 0];
    const int64_t NP = q.n_cite;
    const int64_t P = q.new_off[y] + i;               // sequence-wide citing index
    const int c = i < Dy ? q.cnt[P] : 0;
    const int j0 = lg * K;                            // first slot of this lane
    float wr[K], Lr[K], ar[K];
    int32_t pos[K];  // each slot's by-target position, loaded with its target
    uint32_t evm = 0, vm = 0;
    double ll = 0.0, gaacc = 0.0, gtacc = 0.0;
    if (c > 0) {
#pragma unroll
        for (int e = 0; e < K; ++e) {
            const int j = j0 + e;
            const bool v = j < c;
            const int64_t qi = (int64_t)(v ? j : 0) * NP + P;
            if (GRAD) pos[e] = q.pos_t[qi];
            const int32_t t = q.cit_t[qi];
            const bool ev = v && q.kind_t[qi] != 1;
            const float2 la = q.la_t[qi];
            float x;
            const float w = lik_weight_la(la.x, la.y, a.logfit[t], s, x);
            wr[e] = v ? w : 0.0f;
            Lr[e] = la.x;
            ar[e] = la.y;
            evm |= uint32_t(ev) << e;
            vm |= uint32_t(v) << e;
            ll += ev ? (double)x : 0.0;
            if (GRAD) {
                gaacc += ev ? (double)la.x : 0.0;
                gtacc += ev ? (double)la.y : 0.0;
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < K; ++e) {
            wr[e] = 0.0f;
            Lr[e] = 0.0f;
            ar[e] = 0.0f;
        }
    }
    // prefix of the weights in slot order, chained across the lanes
    double Pw[K];
    double start = 0.0;
#pragma unroll
    for (int r = 0; r < G; ++r) {
        double p = start;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            if (lg == r) Pw[e] = p;
            p += (double)wr[e];
        }
        if (G > 1) {
            const double nxt = __shfl(p, (threadIdx.x & 63) - lg + r, 64);
            if (lg == r + 1) start = nxt;
        }
    }
    double inv[K];
    double tot = 0.0;
#pragma unroll
    for (int e = 0; e < K; ++e) {
        const bool ev = (evm >> e) & 1;
        const double Dv = Z - Pw[e];
        // D is a difference of sums of FP32 values: 0, negative (never for a
        // drawn pick) or at least 2^-149, a normal double, where det_log_pos
        // is det_log; the other two cases as det_log gives them.  Computed
        // for every slot and selected: no branch.
        const double lgD = Dv > 0.0 ? det_log_pos(Dv)
                                    : (Dv == 0.0 ? -__builtin_huge_val() : __builtin_nan(""));
        ll -= ev ? lgD : 0.0;
        if (GRAD) {
            const double rD = 1.0 / Dv;
            inv[e] = ev ? rD : 0.0;
        } else {
            inv[e] = 0.0;
        }
        tot += inv[e];
    }
    double Aacc = 0.0;
    if (GRAD) {
        Aacc = tot;
        // sum of tot over the later lanes of the paper
        double suf = tot;
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const double u = __shfl_down(suf, d, G);
            if (lg + d < G) suf += u;
        }
        double later = suf - tot;
        double sL = 0.0, sa = 0.0;
#pragma unroll
        for (int e = K - 1; e >= 0; --e) {
            const double wl = (double)wr[e] * later;
            sL += wl * (double)Lr[e];
            sa += wl * (double)ar[e];
            if ((vm >> e) & 1) share[pos[e]] = (double)((evm >> e) & 1) + wl;
            later += inv[e];
        }
        gaacc += sL;
        gtacc = (gtacc + sa) * s.itau2;
    }
    const double tl = block_sum_d(ll, red);
    if (threadIdx.x == 0) part[blockIdx.x] = tl;
    if (GRAD) {
        const double tA = block_sum_d(Aacc, red);
        const double tg = block_sum_d(gaacc, red);
        const double tt = block_sum_d(gtacc, red);
        if (threadIdx.x == 0) {
            part[nblk + blockIdx.x] = tA;
            part[2 * nblk + blockIdx.x] = tg;
            part[3 * nblk + blockIdx.x] = tt;
        }
    }
}

// Per year y: the slot blocks [year_blk[y], year_blk[y+1]) summed in order,
// columns 0..ncol-1 of the slot partials -> year_sums[y][1 + c] ... stored
// as: col 0 (log L) -> [2], col 1 (A) -> [1], col 2 -> [3], col 3 -> [4].
__global__ __launch_bounds__(kLikBT) void k_lik_year_ev(int64_t nblk, int ncol,
                                                         const int32_t* __restrict__ year_blk,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ year_sums) {
    __shared__ double red[kLikWaves];
    const int y = blockIdx.x;
    const int64_t b0 = year_blk[y], b1 = year_blk[y + 1];
    for (int c = 0; c < ncol; ++c) {
        double v = 0.0;
        for (int64_t b = b0 + threadIdx.x; b < b1; b += kLikBT) v += part[(int64_t)c * nblk + b];
        const double t = block_sum_d(v, red);
        const int dst = c == 0 ? 2 : (c == 1 ? 1 : c + 1);
        if (threadIdx.x == 0) year_sums[y * kLikYearCols + dst] = t;
    }
}

// Sequence preparation: key = target of each filled slot (n_papers for an
// empty slot, so empty slots sort last), value = slot index.
__global__ __launch_bounds__(kLikBT) void k_lik_slot_keys(int64_t T, int m, int64_t n_papers,
                                                          const int32_t* __restrict__ cnt,
                                                          const int32_t* __restrict__ cit,
                                                          uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kLikBT + threadIdx.x;
    if (i >= T) return;
    const int64_t p = i / m;
    const int64_t sl = i - p * m;
    keys[i] = sl < cnt[p] ? (uint32_t)cit[i] : (uint32_t)n_papers;
    vals[i] = (uint32_t)i;
}

// seg_off[j] = first sorted position with key >= j, j in [0, n_papers].
__global__ __launch_bounds__(kLikBT) void k_lik_seg_off(int64_t n_papers, int64_t T,
                                                        const uint32_t* __restrict__ skey,
                                                        int32_t* __restrict__ seg_off) {
    const int64_t j = (int64_t)blockIdx.x * kLikBT + threadIdx.x;
    if (j > n_papers) return;
    int64_t lo = 0, hi = T;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)skey[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    seg_off[j] = (int32_t)lo;
}

// Slot-major copies (slot s of citing paper p at s*NP + p): targets (0 for
// an empty slot), kinds (1 for an empty slot, so it is never a PA slot) and
// the by-target sorted position of every slot (pos_t[s*NP+p] = i where
// perm[i] = p*m + s).
__global__ __launch_bounds__(kLikBT) void k_lik_slot_major(int64_t T, int m, int64_t NP,
                                                           const int32_t* __restrict__ cnt,
                                                           const int32_t* __restrict__ cit,
                                                           const int8_t* __restrict__ kind,
                                                           const uint32_t* __restrict__ perm,
                                                           int32_t* __restrict__ cit_t,
                                                           int8_t* __restrict__ kind_t,
                                                           uint32_t* __restrict__ pos_t) {
    const int64_t i = (int64_t)blockIdx.x * kLikBT + threadIdx.x;
    if (i >= T) return;
    const int64_t p = i / m, sl = i - p * m;
    const bool filled = sl < cnt[p];
    cit_t[sl * NP + p] = filled ? cit[i] : 0;
    kind_t[sl * NP + p] = filled ? kind[i] : (int8_t)1;
    const uint32_t qv = perm[i];
    const int64_t pp = qv / m, ss = qv - pp * m;
    pos_t[ss * NP + pp] = (uint32_t)i;
}

// Per-target sums of the slot shares (stored in by-target order) over sorted
// positions [0, S) in chunks of kLikSegChunk (kLikSegPer consecutive
// positions per thread).  A run of equal keys that starts and ends inside the
// chunk is final: evt[key] = its sum.  A run that continues from the previous
// chunk gives carry[2*chunk] (its part in this chunk), one that continues into
// the next chunk gives carry[2*chunk+1].  The chunk is loaded coalesced into
// LDS and each thread then takes its consecutive positions from there; the
// segmented scan over the threads' trailing runs is a shuffle scan inside each
// wave and a combine over the waves in order (fixed order: the sums are
// reproducible).
__global__ __launch_bounds__(kLikBT) void k_lik_segsum(int64_t S, const uint32_t* __restrict__ skey,
                                                       const double* __restrict__ share,
                                                       double* __restrict__ evt,
                                                       double* __restrict__ carry) {
    __shared__ uint32_t s_key[kLikSegChunk];
    __shared__ double s_v[kLikSegChunk];
    __shared__ double s_wt[kLikWaves];
    __shared__ int s_wf[kLikWaves];
    const int64_t c0 = (int64_t)blockIdx.x * kLikSegChunk;
    const int n = (int)(S - c0 < kLikSegChunk ? S - c0 : kLikSegChunk);
    {
        uint32_t kk[kLikSegPer];
        double vv[kLikSegPer];
#pragma unroll
        for (int k = 0; k < kLikSegPer; ++k) {
            const int i = k * kLikBT + threadIdx.x;
            kk[k] = i < n ? skey[c0 + i] : 0xFFFFFFFFu;
            vv[k] = i < n ? share[c0 + i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kLikSegPer; ++k) {
            s_key[k * kLikBT + threadIdx.x] = kk[k];
            s_v[k * kLikBT + threadIdx.x] = vv[k];
        }
    }
    __syncthreads();
    const int t0 = threadIdx.x * kLikSegPer;
    uint32_t key[kLikSegPer];
    double val[kLikSegPer];
#pragma unroll
    for (int e = 0; e < kLikSegPer; ++e) {
        key[e] = s_key[t0 + e];
        val[e] = s_v[t0 + e];
    }
    const uint32_t kprev = t0 > 0 ? s_key[t0 - 1] : 0xFFFFFFFFu;
    const uint32_t knext = t0 + kLikSegPer < kLikSegChunk ? s_key[t0 + kLikSegPer] : 0xFFFFFFFFu;
    // thread aggregate: sum of the trailing run, and whether a run starts here
    // (the chunk's first position always starts a run for the scan)
    int head_any = t0 == 0 ? 1 : 0;
    double tail = 0.0;
#pragma unroll
    for (int e = 0; e < kLikSegPer; ++e) {
        const uint32_t kp = e ? key[e - 1] : kprev;
        const bool head = (e == 0 && t0 == 0) || key[e] != kp;
        if (head) {
            head_any = 1;
            tail = 0.0;
        }
        tail += val[e];
    }
    // inclusive segmented scan of (head_any, tail): inside the wave...
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v = tail;
    int f = head_any;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double pv = __shfl_up(v, d, 64);
        const int pf = __shfl_up(f, d, 64);
        if (lane >= d && !f) {
            v = pv + v;
            f = pf;
        }
    }
    if (lane == 63) {
        s_wt[wave] = v;
        s_wf[wave] = f;
    }
    __syncthreads();
    // ...then the earlier waves' trailing run, combined in wave order
    double cw = 0.0;
    for (int u = 0; u < wave; ++u) cw = s_wf[u] ? s_wt[u] : cw + s_wt[u];
    const double incl = f ? v : cw + v;
    const double up = __shfl_up(incl, 1, 64);
    double run = lane ? up : cw;  // carry into this thread (0 for thread 0)
    const uint32_t kfirst = s_key[0];
    const bool cont_in = c0 > 0 && skey[c0 - 1] == kfirst;
    const uint32_t klast = s_key[n - 1];
    const int64_t c1 = c0 + n;
    const bool cont_out = c1 < S && skey[c1] == klast;
    const int64_t i0 = c0 + t0;
#pragma unroll
    for (int e = 0; e < kLikSegPer; ++e) {
        const int64_t i = i0 + e;
        if (i >= c1) break;
        const uint32_t kp = e ? key[e - 1] : kprev;
        if ((e == 0 && t0 == 0) || key[e] != kp) run = 0.0;
        run += val[e];
        const uint32_t kn = e + 1 < kLikSegPer ? key[e + 1] : knext;
        if (i + 1 == c1 || kn != key[e]) {            // last position of its run in the chunk
            const bool starts = !(cont_in && key[e] == kfirst);
            const bool ends = !(cont_out && key[e] == klast);
            if (starts && ends) evt[key[e]] = run;
            if (!starts) carry[2 * blockIdx.x] = run;
            if (!ends) carry[2 * blockIdx.x + 1] = run;
        }
    }
}

// One block: log L, dL/dalpha, dL/dtau from the per-year sums and (GRAD)
// the paper-pass partials.
__global__ __launch_bounds__(kLikBT) void k_lik_final(int n_years, int grad, int64_t nblk_p,
                                                       const double* __restrict__ part_p,
                                                       LikArgs a,
                                                       const double* __restrict__ year_sums,
                                                       double* __restrict__ out,
                                                       double* __restrict__ out2) {
    __shared__ double red[kLikWaves];
    const LikScal s = lik_scal(a);
    double ll = 0.0, ga = 0.0, gt = 0.0;
    if (threadIdx.x == 0) {
        for (int y = 0; y < n_years; ++y) {
            ll += year_sums[y * kLikYearCols + 2];
            if (grad) {
                ga += year_sums[y * kLikYearCols + 3];
                gt += year_sums[y * kLikYearCols + 4];
            }
        }
    }
    double sL = 0.0, sA = 0.0;
    if (grad) {
        double v = 0.0, u = 0.0;
        for (int64_t b = threadIdx.x; b < nblk_p; b += kLikBT) {
            v += part_p[b];
            u += part_p[nblk_p + b];
        }
        sL = block_sum_d(v, red);
        sA = block_sum_d(u, red);
    }
    if (threadIdx.x == 0) {
        double r[3] = {ll, ga - sL, gt - sA * s.itau2};
        for (int c = 0; c < 3; ++c) {
            out[c] = r[c];
            if (out2) out2[c] = r[c];
        }
    }
}

template <bool GRAD, int G, int K>
void cite1(hipStream_t st, int64_t nblk, const LikSeq& q, const LikArgs& a,
           const int32_t* ev_block, const double* year_sums, double* share, double* part) {
    k_lik_cite<GRAD, G, K><<<(unsigned)nblk, kLikBT, 0, st>>>(q, a, ev_block, year_sums, share,
                                                               part);
}

template <bool GRAD, int G>
void cite_k(hipStream_t st, int K, int64_t nblk, const LikSeq& q, const LikArgs& a,
            const int32_t* ev_block, const double* year_sums, double* share, double* part) {
    switch (K) {
        case 4: cite1<GRAD, G, 4>(st, nblk, q, a, ev_block, year_sums, share, part); break;
        case 5: cite1<GRAD, G, 5>(st, nblk, q, a, ev_block, year_sums, share, part); break;
        case 6: cite1<GRAD, G, 6>(st, nblk, q, a, ev_block, year_sums, share, part); break;
        case 7: cite1<GRAD, G, 7>(st, nblk, q, a, ev_block, year_sums, share, part); break;
        default: cite1<GRAD, G, 8>(st, nblk, q, a, ev_block, year_sums, share, part); break;
    }
}

template <bool GRAD>
void launch_cite(hipStream_t st, int G, int K, int64_t nblk, const LikSeq& q, const LikArgs& a,
                 const int32_t* ev_block, const double* year_sums, double* share, double* part) {
    switch (G) {
        case 1: cite_k<GRAD, 1>(st, K, nblk, q, a, ev_block, year_sums, share, part); break;
        case 2: cite_k<GRAD, 2>(st, K, nblk, q, a, ev_block, year_sums, share, part); break;
        case 4: cite_k<GRAD, 4>(st, K, nblk, q, a, ev_block, year_sums, share, part); break;
        default: cite_k<GRAD, 8>(st, K, nblk, q, a, ev_block, year_sums, share, part); break;
    }
}

// Sequence preparation: {log(k+1), age} of slot s's target in its year
// (k = the target's in-degree at the start of the year), slot-major like
// cit_t; zero for empty slots.  Fixed by the sequence.
__global__ __launch_bounds__(kLikBT) void k_lik_slot_la(LikSeq q, float2* __restrict__ la_t) {
    const int64_t NP = q.n_cite;
    const int64_t idx = (int64_t)blockIdx.x * kLikBT + threadIdx.x;
    if (idx >= NP * q.m) return;
    const int64_t j = idx / NP, P = idx - j * NP;
    int a = 0, z = q.n_years;  // year y with new_off[y] <= P < new_off[y+1]
    while (z - a > 1) {
        const int mid = (a + z) >> 1;
        if (q.new_off[mid] <= P) a = mid; else z = mid;
    }
    float2 r = make_float2(0.0f, 0.0f);
    if (j < q.cnt[P]) {
        const int32_t t = q.cit_t[idx];
        r.x = det_logf((float)(q.deg[q.deg_off[a] + t] + 1));
        r.y = (float)(q.year_of[a] - q.birth[t]);
    }
    la_t[idx] = r;
}

}  // namespace

__global__ __launch_bounds__(kLikBT) void k_lik_deg_log(const int32_t* __restrict__ deg, int64_t n,
                                                        float* __restrict__ Lpy) {
    const int64_t i = (int64_t)blockIdx.x * kLikBT + threadIdx.x;
    if (i < n) Lpy[i] = det_logf((float)(deg[i] + 1));
}

void launch_lik_deg_log(hipStream_t st, const int32_t* deg, int64_t n, float* Lpy) {
    if (n <= 0) return;
    k_lik_deg_log<<<(unsigned)((n + kLikBT - 1) / kLikBT), kLikBT, 0, st>>>(deg, n, Lpy);
}

void launch_lik_slot_la(hipStream_t st, const LikSeq& q, float2* la_t) {
    const int64_t T = q.n_cite * q.m;
    if (T <= 0) return;
    k_lik_slot_la<<<(unsigned)((T + kLikBT - 1) / kLikBT), kLikBT, 0, st>>>(q, la_t);
}

bool lik_lanes_ok(int m, int lanes) {
    return (lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8) && lanes * 8 >= m;
}

LikCiteShape lik_cite_shape(int m, int lanes) {
    LikCiteShape sh;
    sh.G = 1;
    while (sh.G * 8 < m) sh.G <<= 1;                  // m <= 64: G <= 8
    if (lanes > 0 && lik_lanes_ok(m, lanes)) sh.G = lanes;
    sh.K = (m + sh.G - 1) / sh.G;
    if (sh.K < 4) sh.K = 4;
    return sh;
}

int64_t lik_paper_blocks(int64_t n_papers) {
    return (n_papers + kLikBT * kLikPPT - 1) / (kLikBT * kLikPPT);
}

int64_t lik_seg_chunks(int64_t S) { return (S + kLikSegChunk - 1) / kLikSegChunk; }

void launch_lik_sort_slots(hipStream_t st, int64_t T, int m, int64_t n_papers,
                           const int32_t* cnt, const int32_t* cit, const int8_t* kind,
                           uint32_t* keys, uint32_t* vals, uint32_t* skey, uint32_t* perm,
                           int32_t* cit_t, int8_t* kind_t, uint32_t* pos_t, uint32_t* k_tmp,
                           uint32_t* v_tmp, int key_bits, int64_t* scratch, int32_t* seg_off) {
    if (T > 0) {
        const unsigned g = (unsigned)((T + kLikBT - 1) / kLikBT);
        k_lik_slot_keys<<<g, kLikBT, 0, st>>>(T, m, n_papers, cnt, cit, keys, vals);
        launch_radix_sort_pairs(st, keys, vals, skey, perm, k_tmp, v_tmp, T, key_bits, scratch);
        k_lik_slot_major<<<g, kLikBT, 0, st>>>(T, m, T / m, cnt, cit, kind, perm, cit_t, kind_t,
                                               pos_t);
    }
    k_lik_seg_off<<<(unsigned)((n_papers + 1 + kLikBT - 1) / kLikBT), kLikBT, 0, st>>>(
        n_papers, T, skey, seg_off);
}

void launch_lik_forward_sums(hipStream_t st, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                             double* part_p, double* year_sums) {
    if (nblk_p > 0)
        k_lik_papers<false><<<(unsigned)nblk_p, kLikBT, 0, st>>>(q, a, year_sums, LikSegs{},
                                                                 nullptr, part_p);
    if (q.n_years > 0) k_lik_year_z<<<q.n_years, kLikBT, 0, st>>>(nblk_p, part_p, year_sums);
}

void launch_lik_slots(hipStream_t st, bool grad, LikCiteShape sh, int64_t nblk_e,
                      const int32_t* ev_block, const int32_t* year_blk, const LikSeq& q,
                      const LikArgs& a, double* year_sums, double* share, double* part_e) {
    if (nblk_e > 0) {
        if (grad) launch_cite<true>(st, sh.G, sh.K, nblk_e, q, a, ev_block, year_sums, share, part_e);
        else launch_cite<false>(st, sh.G, sh.K, nblk_e, q, a, ev_block, year_sums, share, part_e);
    }
    if (q.n_years > 0)
        k_lik_year_ev<<<q.n_years, kLikBT, 0, st>>>(nblk_e, grad ? 4 : 1, year_blk, part_e,
                                                     year_sums);
}

void launch_lik_backward(hipStream_t st, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                         const double* year_sums, int64_t S, const uint32_t* skey,
                         const double* share, const LikSegs& sg, double* grad_phi,
                         double* part_p) {
    if (S > 0)
        k_lik_segsum<<<(unsigned)lik_seg_chunks(S), kLikBT, 0, st>>>(S, skey, share, sg.evt,
                                                                      sg.carry);
    if (nblk_p > 0)
        k_lik_papers<true><<<(unsigned)nblk_p, kLikBT, 0, st>>>(q, a, year_sums, sg, grad_phi,
                                                                part_p);
}

void launch_lik_final(hipStream_t st, int n_years, bool grad, int64_t nblk_p,
                      const double* part_p, const LikArgs& a, const double* year_sums,
                      double* out, double* out2) {
    k_lik_final<<<1, kLikBT, 0, st>>>(n_years, grad ? 1 : 0, nblk_p, part_p, a, year_sums, out,
                                      out2);
}

}  // namespace pg
