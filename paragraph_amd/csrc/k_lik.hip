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
// Passes (one evaluation; every reduction has a fixed order, so log L and
// the two scalar gradients are bit-reproducible; dL/dphi takes FP64 atomics
// from the slot pass and may differ in the last bits between runs):
//   k_lik_papers<false>  paper-major: per block and year, the partial Z_y;
//                        zeroes dL/dphi
//   k_lik_year_z         Z_y = fixed-order sum of the block partials
//   k_lik_slots<GRAD,G>  G lanes per citing paper, one lane per slot:
//                        exclusive group scans give P (and PL, Pa); per
//                        block partials of log L (and A, alpha, tau terms);
//                        dL/dphi scatter
//   k_lik_year_ev        per year: log L, A_y and the slot terms
//   k_lik_papers<true>   recomputes w_j(y) (never stored): dL/dphi_j -=
//                        sum_y A_y w_j(y); partials of the A-weighted sums
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

// The simulator's weight arithmetic (pa_weight), keeping L and x.
__device__ inline float lik_weight(int32_t k, int32_t age, float phi, const LikScal& s, float& L,
                                   float& x) {
    L = det_logf((float)(k + 1));
    const float t1 = __builtin_fmaf(-(float)age, s.inv_tau_f, phi);
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

template <int G>
__device__ inline double grp_incl_scan(double v, int lane) {
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const double u = __shfl_up(v, d, G);
        if (lane >= d) v += u;
    }
    return v;
}

template <int G>
__device__ inline double grp_excl_scan(double v, int lane) {
    const double inc = grp_incl_scan<G>(v, lane);
    const double e = __shfl_up(inc, 1, G);
    return lane ? e : 0.0;
}

// Paper-major pass.  One thread per paper j in [0, n_papers); loops over
// the years in which j exists (j < N_y).
//   BWD = false: part[y * nblk + block] = sum of the block's w_j(y); zeroes
//                grad_phi.
//   BWD = true : grad_phi[j] -= sum_y A_y w_j(y); part[q * nblk + block] for
//                q = 0: sum A_y w L, q = 1: sum A_y w age.
template <bool BWD>
__global__ __launch_bounds__(kLikBT) void k_lik_papers(LikSeq q, LikArgs a,
                                                       const double* __restrict__ year_sums,
                                                       double* __restrict__ grad_phi,
                                                       double* __restrict__ part) {
    __shared__ double s_red[kLikWaves][kLikMaxYears];
    __shared__ double s_red2[kLikWaves];
    const LikScal s = lik_scal(a);
    const int64_t nblk = gridDim.x;
    const int64_t j0 = (int64_t)blockIdx.x * kLikBT;
    const int64_t j = j0 + threadIdx.x;
    const bool in = j < q.n_papers;
    const int wave = threadIdx.x >> 6;
    float phi = 0.0f;
    int32_t b = 0;
    if (in) {
        phi = a.logfit[j];
        b = q.birth[j];
        if (!BWD && grad_phi) grad_phi[j] = 0.0;
    }
    double g = 0.0, gL = 0.0, ga = 0.0;
    for (int y = 0; y < q.n_years; ++y) {
        const int64_t Ny = q.n_exist[y];
        if (j0 >= Ny) {                  // block-uniform: no paper of the block exists
            if (!BWD && threadIdx.x == 0) part[(int64_t)y * nblk + blockIdx.x] = 0.0;
            continue;
        }
        double wv = 0.0, Lv = 0.0, av = 0.0;
        if (j < Ny) {
            const int32_t k = q.deg[q.deg_off[y] + j];
            const int32_t age = q.year_of[y] - b;
            float L, x;
            wv = (double)lik_weight(k, age, phi, s, L, x);
            Lv = (double)L;
            av = (double)age;
        }
        if (BWD) {
            const double Ay = year_sums[y * kLikYearCols + 1];
            const double aw = Ay * wv;
            g += aw;
            gL += aw * Lv;
            ga += aw * av;
        } else {
            const double ws = wave_sum_d(wv);
            if ((threadIdx.x & 63) == 0) s_red[wave][y] = ws;
        }
    }
    if (BWD) {
        if (in && grad_phi) grad_phi[j] -= g;
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

// Slot pass.  Each block holds kLikSlotPapers consecutive citing papers of
// one year (ev_block[2*blk] = year, ev_block[2*blk+1] = first paper within
// the year); G lanes per paper, lane l = slot l.  Partials per block, column
// c at part[c * nblk + blk]: 0 log L, 1 A, 2 alpha terms, 3 tau terms.
template <bool GRAD, int G>
__global__ __launch_bounds__(kLikBT) void k_lik_slots(LikSeq q, LikArgs a,
                                                      const int32_t* __restrict__ ev_block,
                                                      const double* __restrict__ year_sums,
                                                      double* __restrict__ grad_phi,
                                                      double* __restrict__ part) {
    __shared__ double red[kLikWaves];
    constexpr int kPer = kLikBT / G;                 // papers per iteration
    constexpr int kIters = kLikSlotPapers / kPer;
    const LikScal s = lik_scal(a);
    const int64_t nblk = gridDim.x;
    const int y = ev_block[2 * blockIdx.x];
    const int64_t first = ev_block[2 * blockIdx.x + 1];
    const int64_t Dy = q.n_new[y];
    const int64_t base = q.new_off[y];
    const int64_t doff = q.deg_off[y];
    const int32_t year = q.year_of[y];
    const double Z = year_sums[y * kLikYearCols + 0];
    const int lane = threadIdx.x % G;
    const int m = q.m;
    double ll = 0.0, Aacc = 0.0, gaacc = 0.0, gtacc = 0.0;
    for (int it = 0; it < kIters; ++it) {
        const int64_t i = first + (int64_t)it * kPer + threadIdx.x / G;
        const bool paper = i < Dy;
        const int32_t c = paper ? q.cnt[base + i] : 0;
        const bool have = lane < c;
        int32_t t = 0;
        int8_t kd = 1;
        double wd = 0.0, Ld = 0.0, ad = 0.0, xd = 0.0;
        if (have) {
            const int64_t slot = (base + i) * m + lane;
            t = q.cit[slot];
            kd = q.kind[slot];
            const int32_t k = q.deg[doff + t];
            const int32_t age = year - q.birth[t];
            float L, x;
            wd = (double)lik_weight(k, age, a.logfit[t], s, L, x);
            Ld = (double)L;
            ad = (double)age;
            xd = (double)x;
        }
        const bool ev = have && kd != 1;
        const double P = grp_excl_scan<G>(wd, lane);
        const double Dv = Z - P;
        if (ev) ll += xd - det_log(Dv);
        if (GRAD) {
            const double PL = grp_excl_scan<G>(wd * Ld, lane);
            const double Pa = grp_excl_scan<G>(wd * ad, lane);
            const double inv = ev ? 1.0 / Dv : 0.0;
            if (ev) {
                Aacc += inv;
                gaacc += Ld + PL * inv;
                gtacc += (ad + Pa * inv) * s.itau2;
            }
            const double inc = grp_incl_scan<G>(inv, lane);
            const double tot = __shfl(inc, G - 1, G);
            const double later = tot - inc;
            if (have) {
                const double gv = (ev ? 1.0 : 0.0) + wd * later;
                unsafeAtomicAdd(grad_phi + t, gv);
            }
        }
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

template <bool GRAD>
void launch_slots(hipStream_t st, int G, int64_t nblk, const LikSeq& q, const LikArgs& a,
                  const int32_t* ev_block, const double* year_sums, double* grad_phi,
                  double* part) {
    const dim3 grid((unsigned)nblk), block(kLikBT);
    switch (G) {
        case 4: k_lik_slots<GRAD, 4><<<grid, block, 0, st>>>(q, a, ev_block, year_sums, grad_phi, part); break;
        case 8: k_lik_slots<GRAD, 8><<<grid, block, 0, st>>>(q, a, ev_block, year_sums, grad_phi, part); break;
        case 16: k_lik_slots<GRAD, 16><<<grid, block, 0, st>>>(q, a, ev_block, year_sums, grad_phi, part); break;
        case 32: k_lik_slots<GRAD, 32><<<grid, block, 0, st>>>(q, a, ev_block, year_sums, grad_phi, part); break;
        default: k_lik_slots<GRAD, 64><<<grid, block, 0, st>>>(q, a, ev_block, year_sums, grad_phi, part); break;
    }
}

}  // namespace

int lik_group_lanes(int m) {
    int G = 4;
    while (G < m && G < 64) G <<= 1;
    return G;
}

int64_t lik_paper_blocks(int64_t n_papers) { return (n_papers + kLikBT - 1) / kLikBT; }

void launch_lik_forward_sums(hipStream_t st, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                             double* part_p, double* year_sums, double* grad_phi) {
    if (nblk_p > 0)
        k_lik_papers<false><<<(unsigned)nblk_p, kLikBT, 0, st>>>(q, a, year_sums, grad_phi, part_p);
    if (q.n_years > 0) k_lik_year_z<<<q.n_years, kLikBT, 0, st>>>(nblk_p, part_p, year_sums);
}

void launch_lik_slots(hipStream_t st, bool grad, int G, int64_t nblk_e, const int32_t* ev_block,
                      const int32_t* year_blk, const LikSeq& q, const LikArgs& a,
                      double* year_sums, double* grad_phi, double* part_e) {
    if (nblk_e > 0) {
        if (grad) launch_slots<true>(st, G, nblk_e, q, a, ev_block, year_sums, grad_phi, part_e);
        else launch_slots<false>(st, G, nblk_e, q, a, ev_block, year_sums, grad_phi, part_e);
    }
    if (q.n_years > 0)
        k_lik_year_ev<<<q.n_years, kLikBT, 0, st>>>(nblk_e, grad ? 4 : 1, year_blk, part_e,
                                                     year_sums);
}

void launch_lik_backward(hipStream_t st, const LikSeq& q, const LikArgs& a, int64_t nblk_p,
                         const double* year_sums, double* grad_phi, double* part_p) {
    if (nblk_p > 0)
        k_lik_papers<true><<<(unsigned)nblk_p, kLikBT, 0, st>>>(q, a, year_sums, grad_phi, part_p);
}

void launch_lik_final(hipStream_t st, int n_years, bool grad, int64_t nblk_p,
                      const double* part_p, const LikArgs& a, const double* year_sums,
                      double* out, double* out2) {
    k_lik_final<<<1, kLikBT, 0, st>>>(n_years, grad ? 1 : 0, nblk_p, part_p, a, year_sums, out,
                                      out2);
}

}  // namespace pg
