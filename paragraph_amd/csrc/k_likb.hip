// k_likb.hip -- the PA log-likelihood and its gradient at B parameter points
// (alpha_b, tau_b, phi_b) in one pass over a recorded citation sequence.  No
// reference counterpart.  Same model, arithmetic and pass structure as
// k_lik.hip; what changes is the layout, so that every stream of the
// sequence (log(k+1) per paper-year, the per-slot target, kind, record and
// sorted position) is read once for all B points and every random access
// moves B contiguous values:
//   phi      float [paper][B]     log-fitness of point b at j*B + b
//   grad     double[paper][B]     dL/dphi
//   share    double[sorted slot][B]  per-slot dL/dphi shares (by target)
//   partials double[...][B], per-year sums double[year][col][B]
// A lane owns one point b; the B lanes of a paper (or slot group) are
// consecutive, b fastest, so the loads of phi and the stores of the shares
// are contiguous runs of B values.  Sums keep a fixed order (point b's sums
// do not depend on the other points), so results are bit-reproducible.
#include <hip/hip_runtime.h>

#include "pg_internal.h"
#include "pg_math.h"

namespace pg {

namespace {

constexpr int kBT = 256;          // threads per block, every pass
constexpr int kWv = kBT / 64;
constexpr int kYChunk = 8;        // paper passes: years per LDS reduction round
constexpr int kPts = 4;           // paper passes: points per thread (at most)
constexpr int kFwdRows = 8;       // forward paper pass: papers per thread

struct PScal {
    float alpha_f, inv_tau_f;
    double itau2;
};

// Point b's alpha and 1/tau as the simulator's weight arithmetic takes them.
__device__ inline PScal point_scal(const double* __restrict__ at, int b) {
    const double alpha = at[2 * b], tau = at[2 * b + 1];
    PScal s;
    s.alpha_f = (float)alpha;
    s.inv_tau_f = (float)(1.0 / tau);
    s.itau2 = 1.0 / (tau * tau);
    return s;
}

// The weight and its exponent x.  FAST = false: the simulator's arithmetic
// (det_expf, the same bits as the weights the sequence was drawn with).
// FAST = true: the hardware exponential, exp(x) = 2^(x*log2(e)) with the
// product's rounding error and log2(e)'s low part folded back in to first
// order (a few ulp of the exact weight).
template <bool FAST>
__device__ inline float weight_la(float L, float age, float phi, const PScal& s, float& x) {
    const float t1 = __builtin_fmaf(-age, s.inv_tau_f, phi);
    x = __builtin_fmaf(s.alpha_f, L, t1);
    if constexpr (FAST) {
        constexpr float kL2E = 1.44269502162933349609375f;       // log2(e) rounded to float
        constexpr float kL2ELo = 1.925963033500011079013348e-8f;  // log2(e) - kL2E
        constexpr float kLn2 = 0.693147180559945309417232f;
        const float y = x * kL2E;
        const float r = __builtin_fmaf(x, kL2E, -y) + x * kL2ELo;
        const float p = __builtin_amdgcn_exp2f(y);
        return __builtin_fmaf(p, r * kLn2, p);
    } else {
        return det_expf(x);
    }
}

// log D for D > 0.  FAST: the exponent times ln 2 plus log2 of the mantissa
// (in [0.5, 1), rounded to float) by the hardware logarithm; absolute error
// about 1e-7 per term.  Otherwise the simulator's det_log_pos.
template <bool FAST>
__device__ inline double log_pos(double D) {
    if constexpr (FAST) {
        constexpr double kLn2d = 0.6931471805599453094172321;
        const double m = __builtin_amdgcn_frexp_mant(D);
        const int e = __builtin_amdgcn_frexp_exp(D);
        const float l2 = __builtin_amdgcn_logf((float)m);
        return __builtin_fma((double)e, kLn2d, (double)l2 * kLn2d);
    } else {
        return det_log_pos(D);
    }
}

// 1/D for D > 0.  FAST: the hardware reciprocal and one Newton step.
template <bool FAST>
__device__ inline double recip_pos(double D) {
    if constexpr (FAST) {
        const double r = __builtin_amdgcn_rcp(D);
        const double e = __builtin_fma(-D, r, 1.0);
        return __builtin_fma(r, e, r);
    } else {
        return 1.0 / D;
    }
}

// Sum over the lanes of a wave that hold the same point (lane % B): the
// result is in every lane, fixed order (butterflies over the offsets >= B).
template <int B>
__device__ inline double wave_sum_b(double v) {
#pragma unroll
    for (int o = 32; o >= B; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the block of the threads with the same point (threadIdx.x % B):
// returned in the threads t < B (for point t); fixed order.
template <int B>
__device__ inline double block_sum_b(double v, double* red /* [kWv * B] */) {
    v = wave_sum_b<B>(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane < B) red[wave * B + lane] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < B) {
#pragma unroll
        for (int w = 0; w < kWv; ++w) t += red[w * B + threadIdx.x];
    }
    return t;
}

template <int B>
constexpr int seg_chunk() { return kLikBSegPer * (kBT / B); }

// Slot shares of paper j for point b: its segment total (0 without slots).
template <int B>
__device__ inline double slot_share_b(int64_t j, int b, const int32_t* __restrict__ seg_off,
                                      const double* __restrict__ evt) {
    return seg_off[j] == seg_off[j + 1] ? 0.0 : evt[j * B + b];
}

// Points per thread of the paper passes: a thread evaluates P points of each
// of its papers, so a paper-year's log(k+1) load, age and mask serve P
// weights; T = B/P threads share a paper.  Rows: papers per thread.  The
// forward pass reduces every year's sums over the wavefront, so it takes
// more papers per thread; the backward pass keeps per-paper gradients in
// registers, so it takes fewer.
template <int B, bool BWD>
constexpr int papers_pts() { return B >= kPts ? kPts : B; }
template <int B, bool BWD>
constexpr int papers_rows() { return BWD ? kLikBPPT / papers_pts<B, BWD>() : kFwdRows; }

// Paper-major passes over tiles of PPB = NR*kBT/T papers: block blk takes
// tiles blk, blk + nblk, ...; in a tile, thread t holds the points
// b0..b0+P-1 (b0 = (t % T)*P) of the papers j = tile*PPB + r*(kBT/T) + t/T
// (r < NR).  Measured on the 30-year sequence at 16 points: 4 points by 8
// papers per thread forward and 4 by 2 backward (424 and 545 us with the
// fast arithmetic) against one point by 8 papers (624 and 731 us), 2 points
// (486 and 594 us) and a 4-paper forward pass (505 us).
//   BWD = false: part[(y*nblk + blk)*B + b] = sum of the block's w_j(y) at b
//                (per-year totals kept in LDS across the block's tiles:
//                dynamic shared memory of n_years*B doubles).
//   BWD = true : grad[j*B + b] = slot shares - sum_y A_y w_j(y);
//                part[(c*nblk + blk)*B + b], c = 0, 1: sums of A w L, A w age.
// Every sum runs in a fixed order (tiles in order, rows of a thread in
// order, lanes by fixed butterflies, waves in order).
template <bool BWD, int B, bool FAST>
__global__ __launch_bounds__(kBT) void k_likb_papers(LikSeq q, const double* __restrict__ at,
                                                     const float* __restrict__ logfit,
                                                     const double* __restrict__ year_sums,
                                                     LikSegs sg, double* __restrict__ grad,
                                                     double* __restrict__ part) {
    constexpr int P = papers_pts<B, BWD>();
    constexpr int T = B / P;                       // threads per paper
    constexpr int R = kBT / T;                     // papers per row
    constexpr int NR = papers_rows<B, BWD>();      // rows (papers) per thread
    constexpr int PPB = NR * R;
    __shared__ double s_red[kYChunk][kWv * B];
    extern __shared__ double s_tot[];              // [n_years * B] (forward pass)
    const int pg = threadIdx.x % T, b0 = pg * P;
    const int pr = threadIdx.x / T;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    PScal s[P];
#pragma unroll
    for (int k = 0; k < P; ++k) s[k] = point_scal(at, b0 + k);
    const int64_t nblk = gridDim.x;
    const int64_t ntile = (q.n_papers + PPB - 1) / PPB;
    if (!BWD)
        for (int k = threadIdx.x; k < q.n_years * B; k += kBT) s_tot[k] = 0.0;
    double gL[P], ga[P];
#pragma unroll
    for (int k = 0; k < P; ++k) gL[k] = ga[k] = 0.0;
    // paper ids stay below 2^31 (32-bit index arithmetic); a paper that does
    // not exist in a year is computed like the others and masked off
    for (int64_t tile = blockIdx.x; tile < ntile; tile += nblk) {
        const int32_t j0 = (int32_t)(tile * PPB);
        float phi[NR][P];
        int32_t bi[NR];
        double g[NR][P];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int32_t j = j0 + r * R + pr;
            const bool in = j < (int32_t)q.n_papers;
#pragma unroll
            for (int k = 0; k < P; ++k) {
                phi[r][k] = in ? logfit[(uint32_t)j * B + b0 + k] : 0.0f;
                g[r][k] = 0.0;
            }
            bi[r] = in ? q.birth[j] : 0;
        }
        for (int yc = 0; yc < q.n_years; yc += kYChunk) {
            const int ny = q.n_years - yc < kYChunk ? q.n_years - yc : kYChunk;
            if (j0 >= q.n_exist[yc + ny - 1]) continue;   // tile-uniform: no paper exists yet
#pragma unroll
            for (int y4 = 0; y4 < kYChunk; y4 += 4) {
                // four years of log(k+1) loads issued before any is used
                float Lv[4][NR];
#pragma unroll
                for (int yy = 0; yy < 4; ++yy) {
                    const int y = yc + y4 + yy;
                    const bool yin = y4 + yy < ny;
                    const int32_t Ny = yin ? (int32_t)q.n_exist[y] : 0;
                    const float* Lrow = Ny > 0 ? q.Lpy + q.deg_off[y] : q.Lpy;
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const int32_t j = j0 + r * R + pr;
                        Lv[yy][r] = Lrow[j < Ny ? j : 0];
                    }
                }
#pragma unroll
                for (int yy = 0; yy < 4; ++yy) {
                    const int yl = y4 + yy;
                    if (yl >= ny) break;
                    const int y = yc + yl;
                    const int32_t Ny = (int32_t)q.n_exist[y];
                    const int32_t year = q.year_of[y];
                    double Ay[P], zacc[P];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        Ay[k] = BWD ? year_sums[(y * kLikYearCols + 1) * B + b0 + k] : 0.0;
                        zacc[k] = 0.0;
                    }
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const int32_t j = j0 + r * R + pr;
                        const bool in = j < Ny;
                        const float L = Lv[yy][r];
                        const int32_t age = year - bi[r];
                        const float agef = (float)age;
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            float x;
                            const float wf = weight_la<FAST>(L, agef, phi[r][k], s[k], x);
                            const double wv = (double)(in ? wf : 0.0f);
                            if (BWD) {
                                const double aw = in ? Ay[k] * wv : 0.0;
                                g[r][k] += aw;
                                gL[k] += aw * (double)L;
                                ga[k] += aw * (double)age;
                            } else {
                                zacc[k] += wv;
                            }
                        }
                    }
                    if (!BWD) {
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            const double ws = wave_sum_b<T>(zacc[k]);
                            if (lane < T) s_red[yl][wave * B + lane * P + k] = ws;
                        }
                    }
                }
            }
            if (!BWD) {
                __syncthreads();
                // the tile's total per (year, point), waves in order, added
                // to the block's running totals
                for (int k = threadIdx.x; k < ny * B; k += kBT) {
                    const int yl = k / B, bb = k % B;
                    double t = 0.0;
#pragma unroll
                    for (int w = 0; w < kWv; ++w) t += s_red[yl][w * B + bb];
                    s_tot[(yc + yl) * B + bb] += t;
                }
                __syncthreads();
            }
        }
        if (BWD) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int32_t j = j0 + r * R + pr;
                if (j < (int32_t)q.n_papers) {
#pragma unroll
                    for (int k = 0; k < P; ++k)
                        grad[(size_t)j * B + b0 + k] =
                            slot_share_b<B>(j, b0 + k, sg.seg_off, sg.evt) - g[r][k];
                }
            }
        }
    }
    if (BWD) {
        __shared__ double red[kWv * B];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const double tL = block_sum_b<T>(gL[k], red);
            const double tA = block_sum_b<T>(ga[k], red);
            if (threadIdx.x < T) {
                part[(0 * nblk + blockIdx.x) * B + threadIdx.x * P + k] = tL;
                part[(1 * nblk + blockIdx.x) * B + threadIdx.x * P + k] = tA;
            }
        }
    } else {
        __syncthreads();
        for (int k = threadIdx.x; k < q.n_years * B; k += kBT)
            part[((int64_t)(k / B) * nblk + blockIdx.x) * B + (k % B)] = s_tot[k];
    }
}

// Z_y at every point: sum over the paper blocks, fixed order.
template <int B>
__global__ __launch_bounds__(kBT) void k_likb_year_z(int64_t nblk, const double* __restrict__ part,
                                                      double* __restrict__ year_sums) {
    __shared__ double red[kWv * B];
    const int y = blockIdx.x;
    const int b = threadIdx.x % B;
    double v = 0.0;
    for (int64_t k = threadIdx.x / B; k < nblk; k += kBT / B)
        v += part[((int64_t)y * nblk + k) * B + b];
    const double t = block_sum_b<B>(v, red);
    if (threadIdx.x < B) year_sums[(y * kLikYearCols + 0) * B + threadIdx.x] = t;
}

// Slot pass: G*B lanes per citing paper (lane gl = lg*B + b), lane (lg, b)
// owning slots [lg*K, lg*K + K) of its paper at point b.  Block blk takes
// kLikBCitePapers consecutive citing papers of year ev_block[2*blk] from
// ev_block[2*blk+1], kBT/(G*B) of them per round.  As k_lik_cite, for every
// point: D = Z_y - P per PA slot (P chained over the lanes in slot order),
// log L, 1/D, `later`, and each slot's share [PA slot] + w_t * later stored
// at share[pos*B + b].  Partials per block, column c at
// part[(c*nblk + blk)*B + b]: 0 log L, 1 A, 2 alpha terms, 3 tau terms
// (rounds in order, then the waves in order).
template <int G, int K, int B, bool FAST>
__global__ __launch_bounds__(kBT) void k_likb_cite(LikSeq q, const double* __restrict__ at,
                                                   const float* __restrict__ logfit,
                                                   const int32_t* __restrict__ ev_block,
                                                   const double* __restrict__ year_sums,
                                                   double* __restrict__ share,
                                                   double* __restrict__ part) {
    static_assert(G * B <= 64, "a paper's lanes must lie in one wavefront");
    constexpr int PPR = kBT / (G * B);            // citing papers per round
    constexpr int ROUNDS = kLikBCitePapers / PPR;
    static_assert(ROUNDS * PPR == kLikBCitePapers, "rounds must tile the block's papers");
    __shared__ double red[4][kWv * B];
    const int gl = threadIdx.x % (G * B);
    const int b = gl % B, lg = gl / B;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const PScal s = point_scal(at, b);
    const int64_t nblk = gridDim.x;
    const int y = ev_block[2 * blockIdx.x];
    const int64_t first = ev_block[2 * blockIdx.x + 1];
    const int32_t Dy = (int32_t)q.n_new[y];
    const double Z = year_sums[(y * kLikYearCols + 0) * B + b];
    const uint32_t NP = (uint32_t)q.n_cite;       // slots of a sequence stay below 2^31
    const uint32_t P0 = (uint32_t)q.new_off[y];
    double ll = 0.0, Aacc = 0.0, gaacc = 0.0, gtacc = 0.0;
    for (int rd = 0; rd < ROUNDS; ++rd) {
        const int32_t i = (int32_t)first + rd * PPR + threadIdx.x / (G * B);
        if ((int32_t)first + rd * PPR >= Dy) break;   // block-uniform: no paper left
        // a lane past the year's papers reads the year's first paper, masked
        const uint32_t P = P0 + (uint32_t)(i < Dy ? i : 0);
        const int c = i < Dy ? q.cnt[P] : 0;
        const int j0 = lg * K;
        float wr[K], Lr[K], ar[K];
        uint32_t pos[K];
        uint32_t evm = 0, vm = 0;
        double ga1 = 0.0, gt1 = 0.0;
        // every slot is loaded (slot 0 of the paper stands in for the empty
        // ones, which exist for every paper) and masked: no branches
#pragma unroll
        for (int e = 0; e < K; ++e) {
            const int j = j0 + e;
            const bool v = j < c;
            const uint32_t qi = (uint32_t)(v ? j : 0) * NP + P;
            pos[e] = q.pos_t[qi];
            const int32_t t = q.cit_t[qi];
            const bool ev = v && q.kind_t[qi] != 1;
            const float2 la = q.la_t[qi];
            float x;
            const float w = weight_la<FAST>(la.x, la.y, logfit[(uint32_t)t * B + b], s, x);
            wr[e] = v ? w : 0.0f;
            Lr[e] = la.x;
            ar[e] = la.y;
            evm |= uint32_t(ev) << e;
            vm |= uint32_t(v) << e;
            ll += ev ? (double)x : 0.0;
            ga1 += ev ? (double)la.x : 0.0;
            gt1 += ev ? (double)la.y : 0.0;
        }
        // prefix of the weights in slot order: the lane's own prefix plus the
        // sum of the paper's earlier lanes at the same point (an inclusive
        // scan of the lane totals at lane stride B); a fixed association
        double Pw[K];
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            Pw[e] = acc;
            acc += (double)wr[e];
        }
        if (G > 1) {
            double incl = acc;
#pragma unroll
            for (int d = 1; d < G; d <<= 1) {
                const double u = __shfl_up(incl, d * B, G * B);
                if (lg >= d) incl += u;
            }
            const double up = __shfl_up(incl, B, G * B);
            const double start = lg > 0 ? up : 0.0;
#pragma unroll
            for (int e = 0; e < K; ++e) Pw[e] += start;
        }
        double inv[K];
        double tot = 0.0;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            const bool ev = (evm >> e) & 1;
            const double Dv = Z - Pw[e];
            // D is a difference of sums of FP32 values: 0, negative (never
            // for a drawn pick) or at least 2^-149, a normal double, where
            // det_log_pos is det_log; the other two cases as det_log gives them
            const double lgD = Dv > 0.0 ? log_pos<FAST>(Dv)
                                        : (Dv == 0.0 ? -__builtin_huge_val() : __builtin_nan(""));
            const double rD = Dv > 0.0 ? recip_pos<FAST>(Dv) : 1.0 / Dv;
            ll -= ev ? lgD : 0.0;
            inv[e] = ev ? rD : 0.0;
            tot += inv[e];
        }
        Aacc += tot;
        // sum of tot over the paper's later lanes of the same point
        double suf = tot;
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const double u = __shfl_down(suf, d * B, G * B);
            if (lg + d < G) suf += u;
        }
        double later = suf - tot;
        double sL = 0.0, sa = 0.0;
#pragma unroll
        for (int e = K - 1; e >= 0; --e) {
            const double wl = (double)wr[e] * later;
            sL += wl * (double)Lr[e];
            sa += wl * (double)ar[e];
            if ((vm >> e) & 1) share[(size_t)pos[e] * B + b] = (double)((evm >> e) & 1) + wl;
            later += inv[e];
        }
        gaacc += ga1 + sL;
        gtacc += (gt1 + sa) * s.itau2;
    }
    // one reduction of the four columns: waves, then the waves in order
    const double v0 = wave_sum_b<B>(ll), v1 = wave_sum_b<B>(Aacc);
    const double v2 = wave_sum_b<B>(gaacc), v3 = wave_sum_b<B>(gtacc);
    if (lane < B) {
        red[0][wave * B + lane] = v0;
        red[1][wave * B + lane] = v1;
        red[2][wave * B + lane] = v2;
        red[3][wave * B + lane] = v3;
    }
    __syncthreads();
    if (threadIdx.x < 4 * B) {
        const int c = threadIdx.x / B, bb = threadIdx.x % B;
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWv; ++w) t += red[c][w * B + bb];
        part[((int64_t)c * nblk + blockIdx.x) * B + bb] = t;
    }
}

// Per year: the slot blocks [year_blk[y], year_blk[y+1]) summed in order for
// the four columns -> year_sums: log L [2], A [1], alpha terms [3], tau [4].
template <int B>
__global__ __launch_bounds__(kBT) void k_likb_year_ev(int64_t nblk, const int32_t* __restrict__ year_blk,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ year_sums) {
    __shared__ double red[kWv * B];
    const int y = blockIdx.x;
    const int b = threadIdx.x % B;
    const int64_t b0 = year_blk[y], b1 = year_blk[y + 1];
    for (int c = 0; c < 4; ++c) {
        double v = 0.0;
        for (int64_t k = b0 + threadIdx.x / B; k < b1; k += kBT / B)
            v += part[((int64_t)c * nblk + k) * B + b];
        const double t = block_sum_b<B>(v, red);
        const int dst = c == 0 ? 2 : (c == 1 ? 1 : c + 1);
        if (threadIdx.x < B) year_sums[(y * kLikYearCols + dst) * B + threadIdx.x] = t;
    }
}

// Per-target sums of the shares over the sorted positions, one block per
// job; jobs are cut at paper boundaries (built once per sequence), so no run
// crosses a job.
//   job_tag < 0: the papers of positions [job_start, job_start+1), at most
//     seg_chunk<B>() positions; lane group pr (B lanes, one per point) takes
//     kLikBSegPer consecutive positions, and the segmented scan over the
//     groups (shuffles at stride B, then the waves in order) gives each run's
//     total: evt[key*B + b].
//   job_tag >= 0: a piece of at most kLikBBigPiece positions of one paper with
//     more than seg_chunk<B>() positions: its total goes to
//     bigpart[tag*B + b] (groups strided over the piece, then the groups in
//     a fixed order), which k_likb_bigsum adds up per paper.
template <int B>
__global__ __launch_bounds__(kBT) void k_likb_segjob(const int64_t* __restrict__ job_start,
                                                      const int32_t* __restrict__ job_tag,
                                                      const uint32_t* __restrict__ skey,
                                                      const double* __restrict__ share,
                                                      double* __restrict__ evt,
                                                      double* __restrict__ bigpart) {
    constexpr int NG = kBT / B;                    // lane groups per block
    constexpr int GPW = 64 / B;                    // lane groups per wave
    __shared__ double s_wt[kWv][B];
    __shared__ int s_wf[kWv];
    const int b = threadIdx.x % B, pr = threadIdx.x / B;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gw = lane / B;                       // group within the wave
    const int64_t c0 = job_start[blockIdx.x];
    const int n = (int)(job_start[blockIdx.x + 1] - c0);
    const int tag = job_tag[blockIdx.x];
    if (tag >= 0) {                                // block-uniform
        // positions pr, pr + NG, ... added in order, eight loads in flight
        double v = 0.0;
        for (int k0 = 0; k0 < n; k0 += 8 * NG) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u * NG + pr;
                x[u] = k < n ? share[(c0 + k) * B + b] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) v += x[u];
        }
        v = wave_sum_b<B>(v);
        if (lane < B) s_wt[wave][lane] = v;
        __syncthreads();
        if (threadIdx.x < B) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kWv; ++w) t += s_wt[w][threadIdx.x];
            bigpart[(int64_t)tag * B + threadIdx.x] = t;
        }
        return;
    }
    const int t0 = pr * kLikBSegPer;
    uint32_t key[kLikBSegPer];
    double val[kLikBSegPer];
#pragma unroll
    for (int e = 0; e < kLikBSegPer; ++e) {
        const int k = t0 + e;
        key[e] = k < n ? skey[c0 + k] : 0xFFFFFFFFu;
        val[e] = k < n ? share[(c0 + k) * B + b] : 0.0;
    }
    // (a group past the job's end holds no position: no neighbour reads)
    const uint32_t kprev = t0 > 0 && t0 <= n ? skey[c0 + t0 - 1] : 0xFFFFFFFFu;
    const uint32_t knext = t0 + kLikBSegPer < n ? skey[c0 + t0 + kLikBSegPer] : 0xFFFFFFFFu;
    int head_any = t0 == 0 ? 1 : 0;
    double tail = 0.0;
#pragma unroll
    for (int e = 0; e < kLikBSegPer; ++e) {
        const uint32_t kp = e ? key[e - 1] : kprev;
        const bool head = (e == 0 && t0 == 0) || key[e] != kp;
        if (head) {
            head_any = 1;
            tail = 0.0;
        }
        tail += val[e];
    }
    // inclusive segmented scan of (head_any, tail) over the groups of the
    // wave (stride B), then over the waves in order
    double v = tail;
    int f = head_any;
#pragma unroll
    for (int d = 1; d < GPW; d <<= 1) {
        const double pv = __shfl_up(v, d * B, 64);
        const int pf = __shfl_up(f, d * B, 64);
        if (gw >= d && !f) {
            v = pv + v;
            f = pf;
        }
    }
    if (gw == GPW - 1) {
        s_wt[wave][b] = v;
        if (b == 0) s_wf[wave] = f;
    }
    __syncthreads();
    double cw = 0.0;
    for (int u = 0; u < wave; ++u) cw = s_wf[u] ? s_wt[u][b] : cw + s_wt[u][b];
    const double incl = f ? v : cw + v;
    const double up = __shfl_up(incl, B, 64);
    double run = gw ? up : cw;
#pragma unroll
    for (int e = 0; e < kLikBSegPer; ++e) {
        if (t0 + e >= n) break;
        const uint32_t kp = e ? key[e - 1] : kprev;
        if ((e == 0 && t0 == 0) || key[e] != kp) run = 0.0;
        run += val[e];
        const uint32_t kn = e + 1 < kLikBSegPer ? key[e + 1] : knext;
        if (t0 + e + 1 == n || kn != key[e]) evt[(int64_t)key[e] * B + b] = run;
    }
}

// Totals of the papers with more than seg_chunk<B>() positions: thread
// (paper k, point b) adds its pieces in order.
template <int B>
__global__ __launch_bounds__(kBT) void k_likb_bigsum(int32_t n_big, const int32_t* __restrict__ big_paper,
                                                      const int32_t* __restrict__ big_first,
                                                      const double* __restrict__ bigpart,
                                                      double* __restrict__ evt) {
    const int64_t t = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (t >= (int64_t)n_big * B) return;
    const int k = (int)(t / B), b = (int)(t % B);
    const int i0 = big_first[k], i1 = big_first[k + 1];
    double v = 0.0;
    for (int i = i0; i < i1; i += 8) {             // in order, eight loads in flight
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = i + u < i1 ? bigpart[(int64_t)(i + u) * B + b] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) v += x[u];
    }
    evt[(int64_t)big_paper[k] * B + b] = v;
}

// One block: per point, log L, dL/dalpha, dL/dtau into out[3*b + c].
template <int B>
__global__ __launch_bounds__(kBT) void k_likb_final(int n_years, int64_t nblk_p,
                                                     const double* __restrict__ part_p,
                                                     const double* __restrict__ at,
                                                     const double* __restrict__ year_sums,
                                                     double* __restrict__ out,
                                                     double* __restrict__ out2) {
    __shared__ double red[kWv * B];
    const int b = threadIdx.x % B;
    double v = 0.0, u = 0.0;
    for (int64_t k = threadIdx.x / B; k < nblk_p; k += kBT / B) {
        v += part_p[(0 * nblk_p + k) * B + b];
        u += part_p[(1 * nblk_p + k) * B + b];
    }
    const double sL = block_sum_b<B>(v, red);
    const double sA = block_sum_b<B>(u, red);
    if (threadIdx.x < B) {
        const PScal s = point_scal(at, b);
        double ll = 0.0, ga = 0.0, gt = 0.0;
        for (int y = 0; y < n_years; ++y) {
            ll += year_sums[(y * kLikYearCols + 2) * B + b];
            ga += year_sums[(y * kLikYearCols + 3) * B + b];
            gt += year_sums[(y * kLikYearCols + 4) * B + b];
        }
        const double r[3] = {ll, ga - sL, gt - sA * s.itau2};
        for (int c = 0; c < 3; ++c) {
            out[3 * b + c] = r[c];
            if (out2) out2[3 * b + c] = r[c];
        }
    }
}

template <int B, bool FAST>
void run_batch(hipStream_t st, const LikSeq& q, const LikBatchBufs& bb, LikCiteShape sh) {
    const int64_t nP = bb.nblk_p, nE = bb.nblk_e;
    double* part_p = bb.part;
    double* part_e = bb.part + (int64_t)q.n_years * nP * B;
    double* part_b = part_e + 4 * nE * B;
    if (nP > 0)
        k_likb_papers<false, B, FAST><<<(unsigned)nP, kBT, (size_t)q.n_years * B * sizeof(double), st>>>(
            q, bb.at, bb.logfit, bb.year_sums, LikSegs{}, nullptr, part_p);
    if (q.n_years > 0) k_likb_year_z<B><<<q.n_years, kBT, 0, st>>>(nP, part_p, bb.year_sums);
    if (nE > 0) {
#define PG_CITE(GG, KK)                                                                   \
    k_likb_cite<GG, KK, B, FAST><<<(unsigned)nE, kBT, 0, st>>>(q, bb.at, bb.logfit, bb.ev_block, \
                                                         bb.year_sums, bb.share, part_e)
#define PG_CITE_K(GG)                                 \
    switch (sh.K) {                                   \
        case 4: PG_CITE(GG, 4); break;                \
        case 5: PG_CITE(GG, 5); break;                \
        case 6: PG_CITE(GG, 6); break;                \
        case 7: PG_CITE(GG, 7); break;                \
        default: PG_CITE(GG, 8); break;               \
    }
        switch (sh.G) {
            case 1: PG_CITE_K(1); break;
            case 2: PG_CITE_K(2); break;
            case 4: if constexpr (4 * B <= 64) { PG_CITE_K(4); } break;
            default: if constexpr (8 * B <= 64) { PG_CITE_K(8); } break;
        }
#undef PG_CITE_K
#undef PG_CITE
    }
    if (q.n_years > 0)
        k_likb_year_ev<B><<<q.n_years, kBT, 0, st>>>(nE, bb.year_blk, part_e, bb.year_sums);
    if (bb.n_jobs > 0)
        k_likb_segjob<B><<<(unsigned)bb.n_jobs, kBT, 0, st>>>(bb.job_start, bb.job_tag, bb.skey,
                                                               bb.share, bb.evt, bb.bigpart);
    if (bb.n_big > 0)
        k_likb_bigsum<B><<<(unsigned)(((int64_t)bb.n_big * B + kBT - 1) / kBT), kBT, 0, st>>>(
            bb.n_big, bb.big_paper, bb.big_first, bb.bigpart, bb.evt);
    if (nP > 0)
        k_likb_papers<true, B, FAST><<<(unsigned)nP, kBT, 0, st>>>(
            q, bb.at, bb.logfit, bb.year_sums, LikSegs{bb.seg_off, bb.evt, nullptr}, bb.grad,
            part_b);
    k_likb_final<B><<<1, kBT, 0, st>>>(q.n_years, nP, part_b, bb.at, bb.year_sums, bb.out,
                                       bb.out2);
}

}  // namespace

bool lik_batch_ok(int B) { return B == 4 || B == 8 || B == 16; }

int64_t lik_batch_paper_blocks(int64_t n_papers, int B) {
    const int64_t ppb = (int64_t)kLikBPPT * (kBT / B);
    const int64_t tiles = (n_papers + ppb - 1) / ppb;
    return tiles < kLikBPaperBlocks ? tiles : kLikBPaperBlocks;
}

int lik_batch_small_job(int B) { return kLikBSegPer * (kBT / B); }

LikCiteShape lik_batch_cite_shape(int m, int B) {
    LikCiteShape sh = lik_cite_shape(m, 0);
    while (sh.G * B > 64 && sh.K * 2 <= 8) {  // keep a paper's lanes in one wavefront
        sh.G >>= 1;
        sh.K = (m + sh.G - 1) / sh.G;
        if (sh.K < 4) sh.K = 4;
    }
    return sh;
}

bool lik_batch_shape_ok(int m, int B) {
    const LikCiteShape sh = lik_batch_cite_shape(m, B);
    return lik_batch_ok(B) && sh.G * B <= 64 && sh.K <= 8 && sh.G * sh.K >= m;
}

void launch_lik_batch(hipStream_t st, int B, const LikSeq& q, const LikBatchBufs& bb,
                      LikCiteShape sh, bool fast) {
    if (fast) {
        switch (B) {
            case 4: run_batch<4, true>(st, q, bb, sh); break;
            case 8: run_batch<8, true>(st, q, bb, sh); break;
            default: run_batch<16, true>(st, q, bb, sh); break;
        }
        return;
    }
    switch (B) {
        case 4: run_batch<4, false>(st, q, bb, sh); break;
        case 8: run_batch<8, false>(st, q, bb, sh); break;
        default: run_batch<16, false>(st, q, bb, sh); break;
    }
}

}  // namespace pg
