// k_abm.hip -- one year of the reference's scoring model (the agent-based
// model of src/abm.cu, year loop src/kernel.cu:1682-1845) on the device.
//
// The reference computes every score on the host, uploads them, and gives
// each new paper one thread that clears and scans whole-graph bitmaps for
// the generator's 1- and 2-hop neighbourhoods (src/abm.cu:577-690), pushes
// every candidate through a 250-entry heap (src/abm.cu:461-565) and fills
// with rejection-sampled uniform papers (src/abm.cu:386-432).  Here:
//   - scores: two passes over the papers (exact 128-bit integer sums of
//     k^3+1 and f^3+1, a recency histogram), then float4 {P, R, F} per
//     paper, so a candidate costs one 16-byte gather;
//   - the same-year subset: the floor(D*share) smallest Philox keys of the
//     year's papers (a radix sort), a uniform subset of fixed size;
//   - selection: one wavefront per (paper, level) streams the level's
//     candidates (level 2 expands the rows of every level-1 paper, 64 rows
//     at a time, lanes spread over the rows' entries by a prefix search), keys
//     them, and keeps the best t in an LDS buffer that is sorted and
//     deduplicated when it fills: after the first fill only keys below the
//     running threshold are appended, so the buffer work follows t*log(M/t)
//     rather than the M candidates; level 2 drops level-1 members by a binary
//     search of the generator's rows, only for candidates that pass;
//   - rows: one wavefront per paper merges the picks, draws the uniform fill
//     64 attempts at a time (the first new id of each batch in attempt
//     order, the same result as the sequential rejection loop), sorts, and
//     writes the row at its scanned offset for the CSR append.
// Every draw and formula is shared with the host through pg_math.h; the
// oracle restates them independently.
#include <climits>

#include "pg_internal.h"

namespace pg {
namespace {

constexpr int kAT = 256;        // threads of the paper passes
constexpr int kAbmBuf = 512;    // selection buffer (entries) per wavefront

// ---- wave-level helpers ----

// Orders the LDS accesses of one wavefront around this point (the LDS
// executes a wavefront's operations in issue order; the fences keep the
// compiler from moving them across).  Wavefronts of a block never wait for
// each other: each works on its own item with its own trip counts.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ascending bitonic sort of s[0, 64*E) held in LDS, in registers.
template <typename T, int E>
__device__ __forceinline__ void wave_sort(T* s) {
    const int lane = threadIdx.x & 63;
    T v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = s[e * 64 + lane];
#pragma unroll
    for (int K = 2; K <= 64 * E; K <<= 1) {
#pragma unroll
        for (int J = K >> 1; J > 0; J >>= 1) {
            if (J >= 64) {
                const int je = J / 64;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if ((e & je) == 0) {
                        const bool up = ((e * 64 + lane) & K) == 0;
                        const T x = v[e], y = v[e + je];
                        const T lo = x < y ? x : y, hi = x < y ? y : x;
                        v[e] = up ? lo : hi;
                        v[e + je] = up ? hi : lo;
                    }
                }
            } else {
                const bool lower = (lane & J) == 0;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const T o = __shfl_xor(v[e], J, 64);
                    const bool up = ((e * 64 + lane) & K) == 0;
                    const bool keep_min = lower == up;
                    const T lo = v[e] < o ? v[e] : o, hi = v[e] < o ? o : v[e];
                    v[e] = keep_min ? lo : hi;
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) s[e * 64 + lane] = v[e];
}

// Sorts s[0,n) (n <= 64*EMAX) ascending, padding to a power of two >= 64 with pad.
template <typename T, int EMAX>
__device__ void wave_sort_n(T* s, int n, T pad) {
    const int lane = threadIdx.x & 63;
    int P = 64;
    while (P < n) P <<= 1;
    for (int i = n + lane; i < P; i += 64) s[i] = pad;
    wsync();
    if (P == 64) wave_sort<T, 1>(s);
    else if (P == 128) wave_sort<T, 2>(s);
    else if (P == 256) wave_sort<T, 4>(s);
    else if constexpr (EMAX >= 8) wave_sort<T, 8>(s);
    wsync();
}

__device__ __forceinline__ uint64_t lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Sorted s[0,n): keeps the first copy of each value other than ~0 (a
// dropped entry), at most `limit`, in place; returns the count.
__device__ int wave_unique(uint64_t* s, int n, int limit) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = lanes_below();
    int out = 0;
    uint64_t carry = ~0ull;  // no real value equals it
    for (int base = 0; base < n && out < limit; base += 64) {
        const int i = base + lane;
        const uint64_t x = i < n ? s[i] : ~0ull;
        uint64_t prev = __shfl_up(x, 1, 64);
        if (lane == 0) prev = carry;
        const bool k = i < n && x != prev && x != ~0ull;
        const uint64_t mask = __ballot(k);
        const int pos = out + __popcll(mask & below);
        carry = __shfl(x, 63, 64);
        wsync();
        if (k && pos < limit) s[pos] = x;
        wsync();
        out += __popcll(mask);
    }
    return out < limit ? out : limit;
}

// Whether the sorted adj[lo, end) holds x.
__device__ __forceinline__ bool row_has(const int32_t* __restrict__ adj, int32_t lo, int32_t end,
                                        int32_t x) {
    int32_t hi = end;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (adj[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < end && adj[lo] == x;
}

struct AbmSetup {
    uint32_t v;
    int32_t gen, sy;
    AbmPaper p;
    AbmBudget b;
};

__device__ __forceinline__ AbmSetup abm_setup(const AbmYear& y, int64_t l, const int32_t* syt) {
    AbmSetup s;
    s.v = (uint32_t)(y.N + l);
    s.p = abm_paper(y.seed, s.v, y.bag, y.n_bag, y.alpha, y.w_pa, y.w_rec, y.w_fit);
    const u32x4 r = philox(s.v, 0u, y.yi, TAG_ABM_GEN, y.seed);
    s.gen = (int32_t)abm_pick(r.x, r.y, y.N);
    s.sy = syt[l] >= 0 ? 1 : 0;
    s.b = abm_budget(s.p.k, s.sy, s.p.alpha, y.frc);
    return s;
}

// ---- score passes ----

__global__ __launch_bounds__(kAT) void k_abm_fit(int64_t lo, int64_t hi, uint64_t seed,
                                                 const double* __restrict__ cdf,
                                                 int16_t* __restrict__ fit) {
    const int64_t c = lo + (int64_t)blockIdx.x * kAT + threadIdx.x;
    if (c < hi) fit[c] = (int16_t)abm_fitness(seed, (uint32_t)c, cdf);
}

// Per block: 128-bit sum of k^3+1 (lo, hi words), 64-bit sum of f^3+1;
// recency counts added to cnt; flag set if an in-degree reaches 2^21.
__global__ __launch_bounds__(kAT) void k_abm_sums(int64_t N, const int32_t* __restrict__ in_deg,
                                                  const int32_t* __restrict__ birth,
                                                  const int16_t* __restrict__ fit, int32_t year,
                                                  int32_t n_rec, uint64_t* __restrict__ parts,
                                                  int32_t* __restrict__ cnt,
                                                  int32_t* __restrict__ flag) {
    __shared__ int32_t hist[kAbmRecMax];
    __shared__ uint64_t r_lo[kAT], r_hi[kAT], r_f[kAT];
    for (int d = threadIdx.x; d < n_rec; d += kAT) hist[d] = 0;
    __syncthreads();
    uint64_t lo = 0, hi = 0, sf = 0;
    bool big = false;
    for (int64_t c = (int64_t)blockIdx.x * kAT + threadIdx.x; c < N; c += (int64_t)gridDim.x * kAT) {
        const uint64_t k = (uint64_t)(uint32_t)in_deg[c];
        big = big || k >= (1u << 21);
        const uint64_t t = k * k * k + 1;
        lo += t;
        hi += lo < t ? 1 : 0;
        const uint64_t f = (uint64_t)fit[c];
        sf += f * f * f + 1;
        const int64_t d = (int64_t)year - birth[c];
        if (d >= 0 && d < n_rec) atomicAdd(&hist[d], 1);
    }
    if (big) atomicOr(flag, 1);
    r_lo[threadIdx.x] = lo;
    r_hi[threadIdx.x] = hi;
    r_f[threadIdx.x] = sf;
    __syncthreads();
    for (int h = kAT / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint64_t a = r_lo[threadIdx.x], b = r_lo[threadIdx.x + h];
            const uint64_t s = a + b;
            r_hi[threadIdx.x] += r_hi[threadIdx.x + h] + (s < a ? 1 : 0);
            r_lo[threadIdx.x] = s;
            r_f[threadIdx.x] += r_f[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        parts[3 * blockIdx.x] = r_lo[0];
        parts[3 * blockIdx.x + 1] = r_hi[0];
        parts[3 * blockIdx.x + 2] = r_f[0];
    }
    for (int d = threadIdx.x; d < n_rec; d += kAT)
        if (hist[d]) atomicAdd(&cnt[d], hist[d]);
}

// Totals (as FP64: hi*2^64 + lo, and the fitness sum), the sum U of the
// recency probabilities of the differences present (ascending difference),
// and the per-difference recency score p(d) / (count(d) * U).
__global__ __launch_bounds__(kAT) void k_abm_final(int nparts, const uint64_t* __restrict__ parts,
                                                   int32_t n_rec, const double* __restrict__ prob,
                                                   const int32_t* __restrict__ cnt,
                                                   double* __restrict__ tot,
                                                   float* __restrict__ rt) {
    __shared__ uint64_t r_lo[kAT], r_hi[kAT], r_f[kAT];
    __shared__ double s_u;
    uint64_t lo = 0, hi = 0, sf = 0;
    for (int i = threadIdx.x; i < nparts; i += kAT) {
        const uint64_t a = parts[3 * i];
        lo += a;
        hi += parts[3 * i + 1] + (lo < a ? 1 : 0);
        sf += parts[3 * i + 2];
    }
    r_lo[threadIdx.x] = lo;
    r_hi[threadIdx.x] = hi;
    r_f[threadIdx.x] = sf;
    __syncthreads();
    for (int h = kAT / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint64_t a = r_lo[threadIdx.x], b = r_lo[threadIdx.x + h];
            const uint64_t s = a + b;
            r_hi[threadIdx.x] += r_hi[threadIdx.x + h] + (s < a ? 1 : 0);
            r_lo[threadIdx.x] = s;
            r_f[threadIdx.x] += r_f[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot[0] = (double)r_hi[0] * 0x1p64 + (double)r_lo[0];
        tot[1] = (double)r_f[0];
        double U = 0.0;
        for (int d = 0; d < n_rec; ++d)
            if (cnt[d] > 0) U += prob[d];
        s_u = U;
    }
    __syncthreads();
    const double U = s_u;
    for (int d = threadIdx.x; d < n_rec; d += kAT)
        rt[d] = (cnt[d] > 0 && U > 0.0) ? (float)(prob[d] / ((double)cnt[d] * U)) : 0.0f;
}

__global__ __launch_bounds__(kAT) void k_abm_scores(int64_t N, const int32_t* __restrict__ in_deg,
                                                    const int32_t* __restrict__ birth,
                                                    const int16_t* __restrict__ fit, int32_t year,
                                                    int32_t n_rec, const float* __restrict__ rt,
                                                    const double* __restrict__ tot,
                                                    float4* __restrict__ comp) {
    const int64_t c = (int64_t)blockIdx.x * kAT + threadIdx.x;
    if (c >= N) return;
    const uint64_t k = (uint64_t)(uint32_t)in_deg[c];
    const uint64_t f = (uint64_t)fit[c];
    const int64_t d = (int64_t)year - birth[c];
    float4 o;
    o.x = (float)((double)(k * k * k + 1) / tot[0]);
    o.y = (d >= 0 && d < n_rec) ? rt[d] : 0.0f;
    o.z = (float)((double)(f * f * f + 1) / tot[1]);
    o.w = 0.0f;
    comp[c] = o;
}

// ---- same-year subset ----

__global__ __launch_bounds__(kAT) void k_abm_same_keys(int64_t N, int64_t D, uint64_t seed,
                                                       uint32_t yi, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals,
                                                       int32_t* __restrict__ syt) {
    const int64_t l = (int64_t)blockIdx.x * kAT + threadIdx.x;
    if (l >= D) return;
    keys[l] = philox((uint32_t)(N + l), 0u, yi, TAG_ABM_SAME, seed).x;
    vals[l] = (uint32_t)l;
    syt[l] = -1;
}

__global__ __launch_bounds__(kAT) void k_abm_same_mark(int64_t N, int64_t D, int64_t K,
                                                       uint64_t seed, uint32_t yi,
                                                       const uint32_t* __restrict__ order,
                                                       int32_t* __restrict__ syt) {
    const int64_t q = (int64_t)blockIdx.x * kAT + threadIdx.x;
    if (q >= K) return;
    const int64_t l = order[q];
    const u32x4 r = philox((uint32_t)(N + l), 0u, yi, TAG_ABM_SAME, seed);
    const int64_t j = abm_pick(r.z, r.w, D - 1);
    syt[l] = (int32_t)(N + (j < l ? j : j + 1));
}

// ---- selection: one wavefront per (paper, level) ----

// The selection buffer of one wavefront: buf[0, nl) is the current list
// (sorted, distinct, checked), buf[nl, nb) entries appended since.  A
// compaction drops the appended entries of the generator's own rows (level
// 2: checked here, for the few entries that got in, rather than for every
// candidate), sorts, deduplicates and keeps the best t; once t are kept,
// only keys below the t-th enter.
struct SelState {
    uint64_t* buf;
    int nl;          // checked entries at the front
    int nb;          // entries in buf
    int t;           // entries kept
    uint64_t thr;    // keys must be below this to enter (all, until t are kept)
};

__device__ __forceinline__ void sel_compact(SelState& st, bool check, const CsrView& g,
                                            int32_t fa, int32_t fb, int32_t ba, int32_t bb) {
    if (check) {
        for (int i = st.nl + (int)(threadIdx.x & 63); i < st.nb; i += 64) {
            const int32_t c = (int32_t)(uint32_t)st.buf[i];
            if (row_has(g.fwd_adj, fa, fb, c) || row_has(g.bwd_adj, ba, bb, c)) st.buf[i] = ~0ull;
        }
    }
    wave_sort_n<uint64_t, kAbmBuf / 64>(st.buf, st.nb, ~0ull);
    st.nb = wave_unique(st.buf, st.nb, st.t);
    st.nl = st.nb;
    wsync();
    st.thr = (st.t > 0 && st.nb == st.t) ? st.buf[st.t - 1] : ~0ull;
}

constexpr int kSelWaves = 4;   // wavefronts (papers) per block

// One wavefront per new paper: the 1-hop list (best t1 of the generator's
// rows) and then the 2-hop list (best t2 of the rows of every 1-hop entry,
// the generator's own rows excluded), through one LDS buffer.  kc[l] and
// kc[D + l] get the entries keyed per level (with repeats).
__global__ __launch_bounds__(64 * kSelWaves) void k_abm_select(AbmYear y, CsrView g,
                                                   const float4* __restrict__ comp,
                                                   const int32_t* __restrict__ syt,
                                                   int32_t* __restrict__ sel,
                                                   int32_t* __restrict__ nsel,
                                                   int32_t* __restrict__ kc) {
    __shared__ uint64_t bufs[kSelWaves][kAbmBuf];
    const int wid = threadIdx.x >> 6;
    uint64_t* buf = bufs[wid];
    const int64_t l = (int64_t)blockIdx.x * kSelWaves + wid;
    const int lane = threadIdx.x & 63;
    if (l >= y.D) return;
    const AbmSetup s = abm_setup(y, l, syt);
    const int32_t gn = s.gen;
    const int32_t fa = g.fwd_off[gn], fb = g.fwd_off[gn + 1];
    const int32_t ba = g.bwd_off[gn], bb = g.bwd_off[gn + 1];
    const int32_t na = fb - fa, n1 = na + (bb - ba);
    const uint64_t below = lanes_below();
    SelState st{buf, 0, 0, 0, ~0ull};
    bool check = false;
    // key of candidate c (valid lanes other than the generator) and whether
    // it enters: below the threshold at the time of the call
    auto key_of = [&](bool valid, int32_t c, bool& pass) -> uint64_t {
        pass = false;
        uint64_t kv = 0;
        if (valid && c != gn) {
            const float4 q = comp[c];
            kv = abm_key(y.seed, s.v, c, y.yi, q.x, q.y, q.z, s.p.w);
            pass = kv < st.thr;
        }
        return kv;
    };
    // more: candidates follow this batch (an early compaction only pays then)
    auto offer = [&](bool pass, uint64_t kv, bool more) {
        const uint64_t mask = __ballot(pass);
        if (mask) {
            if (pass) buf[st.nb + __popcll(mask & below)] = kv;
            st.nb += __popcll(mask);
            wsync();
            // first compaction as soon as t are held (sets the threshold),
            // later ones when the buffer fills
            if ((more && st.thr == ~0ull && st.nb >= st.t) || st.nb > kAbmBuf - 64)
                sel_compact(st, check, g, fa, fb, ba, bb);
        }
    };
    auto finish = [&](int lev) {
        if (st.t > 0) sel_compact(st, check, g, fa, fb, ba, bb);
        else st.nb = 0;
        int32_t* out = sel + ((int64_t)lev * y.D + l) * kAbmSlots;
        for (int i = lane; i < st.nb; i += 64) out[i] = (int32_t)(uint32_t)buf[i];
        if (lane == 0) nsel[(int64_t)lev * y.D + l] = st.nb;
        wsync();
    };
    // Two batches of 64 candidates per iteration: their loads and keys are
    // independent, so each wavefront keeps twice the memory requests and
    // arithmetic in flight.  A key compared with the threshold of before the
    // first batch's offer may enter needlessly; it is dropped at the next
    // compaction (the threshold only falls).
    auto entry = [&](int32_t p) -> int32_t {
        return p < na ? g.fwd_adj[fa + p] : g.bwd_adj[ba + (p - na)];
    };
    // ---- level 1 ----
    st = SelState{buf, 0, 0, s.b.t1, ~0ull};
    check = false;
    if (st.t > 0) {
        for (int32_t base = 0; base < n1; base += 128) {
            const int32_t p0 = base + lane, p1 = p0 + 64;
            const bool v0 = p0 < n1, v1 = p1 < n1;
            const int32_t c0 = v0 ? entry(p0) : 0, c1 = v1 ? entry(p1) : 0;
            bool s0, s1;
            const uint64_t k0 = key_of(v0, c0, s0), k1 = key_of(v1, c1, s1);
            offer(s0, k0, base + 64 < n1);
            if (base + 64 < n1) offer(s1 && k1 < st.thr, k1, base + 128 < n1);
        }
    }
    finish(0);
    // ---- level 2 ----
    int64_t keyed = 0;  // 2-hop entries streamed (with repeats)
    st = SelState{buf, 0, 0, s.b.t2, ~0ull};
    check = true;
    if (st.t > 0) {
        for (int32_t base = 0; base < n1; base += 64) {
            // 64 level-1 entries: their rows, and the inclusive prefix of the
            // rows' lengths over the lanes
            const int32_t q = base + lane;
            int32_t u = -1, uf = 0, ul = 0, ub = 0;
            int64_t len = 0;
            if (q < n1) {
                u = entry(q);
                if (u != gn) {
                    uf = g.fwd_off[u];
                    ul = g.fwd_off[u + 1] - uf;
                    ub = g.bwd_off[u];
                    len = (int64_t)ul + (g.bwd_off[u + 1] - ub);
                }
            }
            int64_t incl = len;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t x = __shfl_up(incl, o, 64);
                if (lane >= o) incl += x;
            }
            const int64_t total = __shfl(incl, 63, 64);
            keyed += total;
            // the candidate at position P of the chunk's rows
            auto locate = [&](int64_t P) -> int32_t {
                int j = 0, hi = 63;  // first lane whose prefix exceeds P
#pragma unroll
                for (int it = 0; it < 6; ++it) {
                    const int mid = (j + hi) >> 1;
                    if (__shfl(incl, mid, 64) > P) hi = mid; else j = mid + 1;
                }
                const int64_t start = __shfl(incl, j, 64) - __shfl(len, j, 64);
                const int32_t juf = __shfl(uf, j, 64), jul = __shfl(ul, j, 64);
                const int32_t jub = __shfl(ub, j, 64);
                const int32_t r = (int32_t)(P - start);
                if (P >= total) return 0;
                return r < jul ? g.fwd_adj[juf + r] : g.bwd_adj[jub + (r - jul)];
            };
            const bool last_chunk = base + 64 >= n1;
            for (int64_t off = 0; off < total; off += 128) {
                const int64_t P0 = off + lane, P1 = P0 + 64;
                const bool v0 = P0 < total, v1 = P1 < total;
                const int32_t c0 = locate(P0), c1 = locate(P1);
                bool s0, s1;
                const uint64_t k0 = key_of(v0, c0, s0), k1 = key_of(v1, c1, s1);
                offer(s0, k0, !last_chunk || off + 64 < total);
                if (off + 64 < total)
                    offer(s1 && k1 < st.thr, k1, !last_chunk || off + 128 < total);
            }
        }
    }
    finish(1);
    if (lane == 0) {
        kc[l] = s.b.t1 > 0 ? n1 : 0;
        kc[y.D + l] = (int32_t)min(keyed, (int64_t)INT32_MAX);
    }
}

// ---- rows ----

__global__ __launch_bounds__(kAT) void k_abm_count(AbmYear y, const int32_t* __restrict__ syt,
                                                   const int32_t* __restrict__ nsel,
                                                   int32_t* __restrict__ cnt,
                                                   int32_t* __restrict__ info) {
    const int64_t l = (int64_t)blockIdx.x * kAT + threadIdx.x;
    if (l >= y.D) return;
    const AbmSetup s = abm_setup(y, l, syt);
    int32_t p1, p2, nu;
    abm_counts(s.p.k, s.sy, s.b, nsel[l], nsel[y.D + l], y.N, &p1, &p2, &nu);
    cnt[l] = 1 + s.sy + p1 + p2 + nu;
    int32_t* o = info + 6 * l;
    o[0] = s.gen;
    o[1] = s.p.k;
    o[2] = syt[l];
    o[3] = p1;
    o[4] = p2;
    o[5] = nu;
}

__global__ __launch_bounds__(64) void k_abm_write(AbmYear y, const int32_t* __restrict__ syt,
                                                  const int32_t* __restrict__ sel,
                                                  const int32_t* __restrict__ info,
                                                  const int64_t* __restrict__ new_off,
                                                  int64_t fwd_base, int32_t* __restrict__ fwd_off,
                                                  int32_t* __restrict__ fwd_adj,
                                                  int32_t* __restrict__ src_of) {
    __shared__ int32_t row[kAbmSlots];
    const int64_t l = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t v = (uint32_t)(y.N + l);
    const int32_t* in = info + 6 * l;
    const int32_t gen = in[0], sy_t = in[2], p1 = in[3], p2 = in[4], nu = in[5];
    const int32_t* s1 = sel + l * kAbmSlots;
    const int32_t* s2 = sel + (y.D + l) * kAbmSlots;
    int n = 1 + p1 + p2;
    for (int i = lane; i < n; i += 64) row[i] = i == 0 ? gen : (i <= p1 ? s1[i - 1] : s2[i - 1 - p1]);
    wave_sort_n<int32_t, kAbmSlots / 64>(row, n, INT_MAX);
    const uint64_t below = lanes_below();
    int need = nu;
    for (uint32_t a0 = 0; need > 0; a0 += 64) {
        const u32x4 r = philox(v, a0 + (uint32_t)lane, y.yi, TAG_ABM_UNI, y.seed);
        const int32_t c = (int32_t)abm_pick(r.x, r.y, y.N);
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (row[mid] < c) lo = mid + 1; else hi = mid;
        }
        bool ok = !(lo < n && row[lo] == c);
        for (int q = 0; q < 63; ++q) {
            const int32_t o = __shfl(c, q, 64);
            if (q < lane && o == c) ok = false;
        }
        const uint64_t mask = __ballot(ok);
        const int rank = __popcll(mask & below);
        wsync();
        if (ok && rank < need) row[n + rank] = c;
        const int took = min(__popcll(mask), need);
        n += took;
        need -= took;
        wsync();
        if (took) wave_sort_n<int32_t, kAbmSlots / 64>(row, n, INT_MAX);
    }
    if (sy_t >= 0 && lane == 0) row[n] = sy_t;  // a new paper: past every existing id
    n += sy_t >= 0 ? 1 : 0;
    wsync();
    const int64_t b = new_off[l];
    for (int i = lane; i < n; i += 64) {
        fwd_adj[fwd_base + b + i] = row[i];
        src_of[b + i] = (int32_t)v;
    }
    if (lane == 0) fwd_off[y.N + l + 1] = (int32_t)(fwd_base + new_off[l + 1]);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kAT - 1) / kAT); }

}  // namespace

void launch_abm_fit(hipStream_t s, int64_t lo, int64_t hi, uint64_t seed, const double* cdf,
                    int16_t* fit) {
    if (hi > lo) k_abm_fit<<<grid_of(hi - lo), kAT, 0, s>>>(lo, hi, seed, cdf, fit);
}

int abm_sum_parts(int64_t N) {
    const int64_t b = (N + kAT - 1) / kAT;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, kAbmMaxParts));
}

void launch_abm_scores(hipStream_t s, int64_t N, const int32_t* in_deg, const int32_t* birth,
                       const int16_t* fit, int32_t year, int32_t n_rec, const double* prob,
                       uint64_t* parts, int32_t* cnt, int32_t* flag, double* tot, float* rt,
                       float4* comp) {
    const int np = abm_sum_parts(N);
    k_abm_sums<<<np, kAT, 0, s>>>(N, in_deg, birth, fit, year, n_rec, parts, cnt, flag);
    k_abm_final<<<1, kAT, 0, s>>>(np, parts, n_rec, prob, cnt, tot, rt);
    if (N > 0) k_abm_scores<<<grid_of(N), kAT, 0, s>>>(N, in_deg, birth, fit, year, n_rec, rt, tot, comp);
}

void launch_abm_same_keys(hipStream_t s, int64_t N, int64_t D, uint64_t seed, uint32_t yi,
                          uint32_t* keys, uint32_t* vals, int32_t* syt) {
    if (D > 0) k_abm_same_keys<<<grid_of(D), kAT, 0, s>>>(N, D, seed, yi, keys, vals, syt);
}

void launch_abm_same_mark(hipStream_t s, int64_t N, int64_t D, int64_t K, uint64_t seed,
                          uint32_t yi, const uint32_t* order, int32_t* syt) {
    if (K > 0) k_abm_same_mark<<<grid_of(K), kAT, 0, s>>>(N, D, K, seed, yi, order, syt);
}

void launch_abm_select(hipStream_t s, const AbmYear& y, const CsrView& g, const float4* comp,
                       const int32_t* syt, int32_t* sel, int32_t* nsel, int32_t* kc) {
    if (y.D > 0)
        k_abm_select<<<(uint32_t)((y.D + kSelWaves - 1) / kSelWaves), 64 * kSelWaves, 0, s>>>(
            y, g, comp, syt, sel, nsel, kc);
}

void launch_abm_count(hipStream_t s, const AbmYear& y, const int32_t* syt, const int32_t* nsel,
                      int32_t* cnt, int32_t* info) {
    if (y.D > 0) k_abm_count<<<grid_of(y.D), kAT, 0, s>>>(y, syt, nsel, cnt, info);
}

void launch_abm_write(hipStream_t s, const AbmYear& y, const int32_t* syt, const int32_t* sel,
                      const int32_t* info, const int64_t* new_off, int64_t fwd_base,
                      int32_t* fwd_off, int32_t* fwd_adj, int32_t* src_of) {
    if (y.D > 0)
        k_abm_write<<<(uint32_t)y.D, 64, 0, s>>>(y, syt, sel, info, new_off, fwd_base, fwd_off,
                                                 fwd_adj, src_of);
}

}  // namespace pg
