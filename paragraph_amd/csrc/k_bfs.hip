// k_bfs.hip -- 2-hop neighbourhood of a root paper and the neighbourhood
// draws of PA (phases 2 and 3 of a new paper).
//
// Replaces Stage 1 of the reference (kernelCallStage1 / _warped,
// src/kernel.cu:17-168, bodies GetOneAndTwoHopNeighborhood[_Warp],
// src/abm.cu:577-883), which cleared and scanned three whole-graph bitmaps
// per query (Theta(N) per paper), plus the neighbourhood sampling of
// Stage 2/3 (src/abm.cu:887-956) and the shortfall fill of Stage 4
// (src/abm.cu:958-994).
//
// Semantics (identical to the oracle): F1 = (fwd(root) U bwd(root)) \ {root};
// F1c = its `cap` smallest ids; F2 = (U_{v in F1c} fwd(v) U bwd(v)) \ F1 \
// {root}; S = the `cap` smallest ids of F1c U F2, ascending.
//
// One 256-thread workgroup per query.  Per-query work is proportional to the
// adjacency it touches, not to N: the candidate ids of each level are
// gathered into a 4096-entry LDS window, bitonic-sorted and deduplicated in
// LDS.  When a level has more than 4096 candidate entries, the id range is
// cut into windows [lo,hi) chosen by per-row binary search so that each
// window fits, and windows are processed in ascending order until the cap is
// reached.  Rows of the CSR are sorted, so every row's part in a window is a
// contiguous slice.
#include <climits>

#include "pg_internal.h"

namespace pg {
namespace {

constexpr int kBT = 256;             // threads per query
constexpr int kWin = 4096;           // LDS window entries

struct BfsLds {
    // big[0, kBfsCapMax): 1-hop list; big[kBfsCapMax, 2*kBfsCapMax): window.
    // After the BFS the whole array is reused as double[kBfsCapMax] (the
    // neighbourhood CDF) with the gathered FP32 weights in its upper half.
    int32_t big[2 * kBfsCapMax];
    int32_t S[kBfsCapMax];
    int32_t cited[kMaxCites];
    int64_t red[kBT / 64];
    int32_t iscan[kBT / 64];
    int64_t tot[3];      // candidate entries per level (first window)
    int32_t nout1;       // |F1c| of the current query
    int32_t fill;
    int32_t lo, hi, nout;
    int32_t flag;
};

// One LDS object per workgroup; every helper below addresses it directly so
// the compiler keeps LDS address-space accesses (ds_*) throughout.
__shared__ BfsLds L;

__device__ __forceinline__ int lower_bound_i32(const int32_t* p, int b, int e, int32_t x) {
    while (b < e) {
        int mid = (b + e) >> 1;
        if (p[mid] < x) b = mid + 1; else e = mid;
    }
    return b;
}

__device__ __forceinline__ bool row_contains(const int32_t* p, int b, int e, int32_t x) {
    int k = lower_bound_i32(p, b, e, x);
    return k < e && p[k] == x;
}

__device__ int64_t block_sum_i64(int64_t v) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if (lane == 0) L.red[wid] = v;
    __syncthreads();
    int64_t t = 0;
    for (int k = 0; k < kBT / 64; ++k) t += L.red[k];
    return t;
}

// Exclusive scan of one int per thread (block-wide); *total gets the sum.
__device__ int block_excl_scan_i32(int v, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) L.iscan[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int k = 0; k < kBT / 64; ++k) { if (k < wid) pre += L.iscan[k]; tot += L.iscan[k]; }
    *total = tot;
    return pre + x - v;
}

__device__ void bitonic_sort(int32_t* s, int P) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += kBT) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int32_t a = s[i], b = s[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) { s[i] = b; s[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
}

struct Query {
    CsrView g;
    int32_t root;
    int level;          // 1: rows of the root; 2: F1c list + rows of F1c
    int n1c;            // |F1c| (level 2)
    bool check_f1;      // level 2: F1 was capped, so filter F1 \ F1c
    int32_t f1max;
};

__device__ __forceinline__ int num_rows(const Query& q) { return q.level == 1 ? 2 : 1 + 2 * q.n1c; }

__device__ __forceinline__ void row_of(const Query& q, const int32_t* f1, int r, const int32_t*& p,
                                       int& b, int& e) {
    int32_t v;
    int d;
    if (q.level == 1) {
        v = q.root; d = r;
    } else if (r == 0) {
        p = f1; b = 0; e = q.n1c; return;
    } else {
        v = f1[(r - 1) >> 1]; d = (r - 1) & 1;
    }
    const int32_t* off = d ? q.g.bwd_off : q.g.fwd_off;
    p = d ? q.g.bwd_adj : q.g.fwd_adj;
    b = off[v];
    e = off[v + 1];
}

// Number of row entries with id in [lo, hi) (hi == INT_MAX: no upper bound).
__device__ int64_t count_range(const Query& q, const int32_t* f1, int32_t lo, int32_t hi) {
    int64_t acc = 0;
    const int nr = num_rows(q);
    for (int r = threadIdx.x; r < nr; r += kBT) {
        const int32_t* p; int b, e;
        row_of(q, f1, r, p, b, e);
        int a = lo > 0 ? lower_bound_i32(p, b, e, lo) : b;
        int z = hi != INT_MAX ? lower_bound_i32(p, a, e, hi) : e;
        acc += z - a;
    }
    return block_sum_i64(acc);
}

__device__ __forceinline__ bool keep_id(const Query& q, int32_t x) {
    if (x == q.root) return false;
    if (q.level == 2 && q.check_f1 && x > q.f1max) {
        const int32_t* fo = q.g.fwd_off; const int32_t* bo = q.g.bwd_off;
        if (row_contains(q.g.fwd_adj, fo[q.root], fo[q.root + 1], x)) return false;
        if (row_contains(q.g.bwd_adj, bo[q.root], bo[q.root + 1], x)) return false;
    }
    return true;
}

// Collects the `cap` smallest distinct kept ids of the level's rows into
// L.S (to_s) or the 1-hop list L.big (ascending).  Kept out of line: inlining
// it into both kernels crashes the ROCm 7.2 optimiser.  Returns the count.  All threads of the block call it.
__device__ __attribute__((noinline)) int collect(const Query& q, bool to_s, int cap) {
    const int32_t n = (int32_t)q.g.n;
    const int32_t* f1 = L.big;
    int32_t* win = L.big + kBfsCapMax;
    int32_t* out = to_s ? L.S : L.big;
    if (threadIdx.x == 0) { L.lo = 0; L.nout = 0; L.tot[q.level] = 0; }
    __syncthreads();
    for (;;) {
        const int32_t lo = L.lo;
        const int64_t rem = count_range(q, f1, lo, INT_MAX);
        if (lo == 0 && threadIdx.x == 0) L.tot[q.level] = rem;
        if (rem == 0) break;
        int32_t hi;
        if (rem <= kWin) {
            hi = INT_MAX;
        } else {
            int32_t a = lo + 1, z = n;  // largest a with count(lo, a) <= kWin
            if (count_range(q, f1, lo, a) > kWin) {
                hi = a;                  // single id window (truncated below)
            } else {
                while (z - a > 1) {
                    int32_t mid = a + (z - a) / 2;
                    if (count_range(q, f1, lo, mid) <= kWin) a = mid; else z = mid;
                }
                hi = a;
            }
        }
        // gather [lo, hi) into the window
        if (threadIdx.x == 0) L.fill = 0;
        __syncthreads();
        {
            const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
            const int nr = num_rows(q);
            for (int r = wid; r < nr; r += kBT / 64) {
                const int32_t* p; int b, e;
                row_of(q, f1, r, p, b, e);
                int a = lo > 0 ? lower_bound_i32(p, b, e, lo) : b;
                int z = hi != INT_MAX ? lower_bound_i32(p, a, e, hi) : e;
                int len = z - a;
                if (len <= 0) continue;
                int base = 0;
                if (lane == 0) base = atomicAdd(&L.fill, len);
                base = __shfl(base, 0, 64);
                for (int k = lane; k < len; k += 64) {
                    int pos = base + k;
                    if (pos < kWin) win[pos] = p[a + k];
                }
            }
        }
        __syncthreads();
        const int nw = L.fill < kWin ? L.fill : kWin;
        int P = 1;
        while (P < nw) P <<= 1;
        for (int k = nw + threadIdx.x; k < P; k += kBT) win[k] = INT_MAX;
        __syncthreads();
        if (nw > 1) bitonic_sort(win, P);
        // unique + filter, ordered compaction
        const int chunk = (nw + kBT - 1) / kBT;
        const int s0 = threadIdx.x * chunk;
        const int s1 = min(nw, s0 + chunk);
        int mine = 0;
        for (int k = s0; k < s1; ++k) {
            int32_t x = win[k];
            if ((k == 0 || win[k - 1] != x) && keep_id(q, x)) ++mine;
        }
        int tot;
        int pos = block_excl_scan_i32(mine, &tot) + L.nout;
        for (int k = s0; k < s1; ++k) {
            int32_t x = win[k];
            if ((k == 0 || win[k - 1] != x) && keep_id(q, x)) {
                if (pos < cap) out[pos] = x;
                ++pos;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int no = L.nout + tot;
            L.nout = no < cap ? no : cap;
            L.lo = hi;
        }
        __syncthreads();
        if (L.nout >= cap || hi == INT_MAX || hi >= n) break;
    }
    __syncthreads();
    return L.nout;
}

// Builds S for `root` into L.S; returns |S|.
__device__ int two_hop(const CsrView& g, int32_t root, int cap) {
    int32_t* f1 = L.big;
    Query q{g, root, 1, 0, false, 0};
    const int n1c = collect(q, false, cap);
    if (threadIdx.x == 0) L.nout1 = n1c;
    q.level = 2;
    q.n1c = n1c;
    q.check_f1 = false;
    q.f1max = 0;
    if (n1c == cap) {
        q.f1max = f1[cap - 1];
        const int32_t* fo = g.fwd_off; const int32_t* bo = g.bwd_off;
        int32_t mx = -1;
        if (fo[root + 1] > fo[root]) mx = max(mx, g.fwd_adj[fo[root + 1] - 1]);
        if (bo[root + 1] > bo[root]) mx = max(mx, g.bwd_adj[bo[root + 1] - 1]);
        q.check_f1 = mx > q.f1max;
    }
    return collect(q, true, cap);
}

__device__ __forceinline__ bool in_cited(const int32_t* c, int n, int32_t x) {
    for (int j = 0; j < n; ++j) if (c[j] == x) return true;
    return false;
}

__global__ __launch_bounds__(kBT) void k_bfs_nb(CsrView g, const float* __restrict__ w,
                                                const double* __restrict__ cdf, int m, int n_nb,
                                                int cap, uint64_t seed, uint32_t yi,
                                                int32_t* __restrict__ cit, int8_t* __restrict__ kind,
                                                int32_t* __restrict__ cnt, int32_t* __restrict__ near,
                                                int32_t* __restrict__ anext,
                                                unsigned long long* __restrict__ bytes_out) {
    const int64_t i = blockIdx.x;
    const int64_t N = g.n;
    const uint32_t gid = (uint32_t)(N + i);
    const int c0 = cnt[i];
    if (c0 <= 0 || n_nb <= 0) return;  // uniform across the block
    for (int j = threadIdx.x; j < c0; j += kBT) L.cited[j] = cit[i * m + j];
    __syncthreads();
    const int ns = two_hop(g, L.cited[0], cap);
    if (threadIdx.x == 0) {
        // algorithmic bytes: offsets + adjacency of the root and of each
        // capped 1-hop paper, plus the output ids
        const int64_t b = 16 * (1 + (int64_t)L.nout1) + 4 * (L.tot[1] + L.tot[2] - L.nout1) +
                          4 * (int64_t)ns;
        atomicAdd(bytes_out, (unsigned long long)b);
    }

    // available = S \ cited
    int mine = 0;
    for (int s = threadIdx.x; s < ns; s += kBT) mine += in_cited(L.cited, c0, L.S[s]) ? 0 : 1;
    const int64_t avail = block_sum_i64(mine);
    int c = c0;
    if (avail <= n_nb) {
        // take every available neighbour, ascending
        const int chunk = (ns + kBT - 1) / kBT;
        const int s0 = threadIdx.x * chunk, s1 = min(ns, s0 + chunk);
        int cntm = 0;
        for (int s = s0; s < s1; ++s) cntm += in_cited(L.cited, c0, L.S[s]) ? 0 : 1;
        int tot;
        int pos = block_excl_scan_i32(cntm, &tot) + c0;
        for (int s = s0; s < s1; ++s) {
            int32_t x = L.S[s];
            if (!in_cited(L.cited, c0, x)) {
                cit[i * m + pos] = x;
                kind[i * m + pos] = 1;
                ++pos;
            }
        }
        __syncthreads();
        for (int s = s0, p2 = pos - cntm; s < s1; ++s) {
            int32_t x = L.S[s];
            if (!in_cited(L.cited, c0, x)) L.cited[p2++] = x;
        }
        __syncthreads();
        c = c0 + tot;
    } else {
        float* wv = reinterpret_cast<float*>(L.big + kBfsCapMax);
        double* D = reinterpret_cast<double*>(L.big);
        for (int s = threadIdx.x; s < ns; s += kBT) wv[s] = w[L.S[s]];
        __syncthreads();
        if (threadIdx.x == 0) {
            // Sequential FP64 prefix in ascending id order (in place: D[s]
            // only overwrites weights that have already been read).
            double acc = 0.0;
            for (int s = 0; s < ns; ++s) { acc += (double)wv[s]; D[s] = acc; }
            const double Zs = acc;
            const int b_max = kAttemptsPerCite * n_nb;
            int got = 0;
            for (int bt = 0; got < n_nb && bt < b_max; ++bt) {
                u32x4 r = philox(gid, (uint32_t)bt, yi, TAG_NB, seed);
                double t = u53(r.x, r.y) * Zs;
                int lo = 0, hi = ns;
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (D[mid] > t) hi = mid; else lo = mid + 1;
                }
                if (lo >= ns) lo = ns - 1;
                int32_t x = L.S[lo];
                if (!in_cited(L.cited, c, x)) {
                    L.cited[c] = x;
                    cit[i * m + c] = x;
                    kind[i * m + c] = 1;
                    ++c;
                    ++got;
                }
            }
            L.fill = c;
        }
        __syncthreads();
        c = L.fill;
    }
    if (threadIdx.x == 0) {
        // refill with PA draws, continuing the attempt sequence
        const double Z = cdf[N - 1];
        const double eps = kNearRel * Z;
        const int a_max = kAttemptsPerCite * m;
        int a = anext[i];
        int nr = 0;
        while (c < m && a < a_max) {
            u32x4 r = philox(gid, (uint32_t)a, yi, TAG_PA, seed);
            double t = u53(r.x, r.y) * Z;
            int64_t k = cdf_upper_bound(cdf, N, t);
            if (cdf_near(cdf, N, k, t, eps)) ++nr;
            ++a;
            if (!in_cited(L.cited, c, (int32_t)k)) {
                L.cited[c] = (int32_t)k;
                cit[i * m + c] = (int32_t)k;
                kind[i * m + c] = 2;
                ++c;
            }
        }
        cnt[i] = c;
        near[i] += nr;
        anext[i] = a;
    }
}

__global__ __launch_bounds__(kBT) void k_bfs_only(CsrView g, const int32_t* __restrict__ roots,
                                                  int cap, int32_t* __restrict__ out,
                                                  int32_t* __restrict__ out_len) {
    const int64_t i = blockIdx.x;
    const int32_t root = roots[i];
    int ns = 0;
    if (root >= 0 && root < g.n) ns = two_hop(g, root, cap);
    for (int s = threadIdx.x; s < ns; s += kBT) out[i * cap + s] = L.S[s];
    if (threadIdx.x == 0) out_len[i] = ns;
}

}  // namespace

void launch_bfs_nb(hipStream_t s, const CsrView& g, const float* w, const double* cdf, int64_t n_new,
                   int m, int n_nb, int cap, uint64_t seed, uint32_t year_index, int32_t* cit,
                   int8_t* kind, int32_t* cnt, int32_t* near, int32_t* anext,
                   unsigned long long* bytes_out) {
    if (n_new <= 0 || n_nb <= 0) return;
    hipLaunchKernelGGL(k_bfs_nb, dim3((uint32_t)n_new), dim3(kBT), 0, s, g, w, cdf, m, n_nb, cap,
                       seed, year_index, cit, kind, cnt, near, anext, bytes_out);
}

void launch_bfs_only(hipStream_t s, const CsrView& g, int64_t count, const int32_t* roots, int cap,
                     int32_t* out, int32_t* out_len) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_bfs_only, dim3((uint32_t)count), dim3(kBT), 0, s, g, roots, cap, out,
                       out_len);
}

}  // namespace pg
