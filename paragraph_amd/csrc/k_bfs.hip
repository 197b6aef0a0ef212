// k_bfs.hip -- 2-hop neighbourhoods and the neighbourhood draws of PA
// (phases 2 and 3 of a new paper).
//
// Replaces Stage 1 of the reference (kernelCallStage1 / _warped,
// src/kernel.cu:17-168, bodies GetOneAndTwoHopNeighborhood[_Warp],
// src/abm.cu:577-883), which cleared and scanned three whole-graph bitmaps
// per query (Theta(N) per paper), the neighbourhood sampling of Stage 2/3
// (src/abm.cu:887-956) and the shortfall fill of Stage 4 (src/abm.cu:958-994).
//
// Semantics (identical to the oracle): F1 = (fwd(root) U bwd(root)) \ {root};
// F1c = its `cap` smallest ids; F2 = (U_{v in F1c} fwd(v) U bwd(v)) \ F1 \
// {root}; S = the `cap` smallest ids of F1c U F2, ascending.
//
// The neighbourhood depends only on the root and on the graph, which is
// frozen for the year, and PA roots concentrate on hubs (late in a 10-year
// run thousands of papers share one root).  So the year's roots are
// deduplicated (a bitmap over paper ids, numbered in id order), S and its
// CDF are built once per distinct root, and each paper then draws from its
// root's S (k_nb_draw).
//
// S is built by three tiers, each handing the roots over its limits to the
// next: one wavefront per root for small and medium neighbourhoods
// (k_bfs_small, k_bfs_mid: rows gathered into LDS, sorted by a register
// bitonic network, compacted by ballots), then one 512-thread workgroup per
// root (k_bfs_root): the entries below the answer bound are radix-sorted in
// LDS -- up to four rows longer than the cap (level 2) or than one
// wavefront's sort (level 1) are merged into the sorted list instead, rows
// being ascending -- or, when they do not fit, staged in id buckets in a
// device slot or swept through the id range in bitmap windows.  Per-query
// work follows the adjacency touched below the final threshold, not N.
//
// The neighbourhood CDF is exact 64-bit fixed point (weights scaled so the
// largest maps below 2^50), built by an integer scan, so it is
// bit-identical to the oracle's sequential loop; a 64-bucket index of it
// starts each draw's search.
#include <climits>

#include "pg_internal.h"

namespace pg {
namespace {

constexpr int kBT = 512;                    // threads per large root
constexpr int kWaves = kBT / 64;
constexpr int kWin = 7168;                  // LDS buffer: sort window or bitmap words
constexpr int kSortMax = 4096;              // totals up to this are always radix-sorted
constexpr int kRadixMax = 7168;             // radix path: entries sorted in LDS (L.win)
constexpr int kRadixRounds = kRadixMax / kBT;
constexpr int kHistBuckets = 1024;          // radix path: buckets of the cut histogram
constexpr int kWaveSortMax = 512;           // radix path: fewer entries sorted by one wavefront
constexpr int kBmBits = kWin * 32;          // ids per bitmap window
constexpr int kMaxRows = 1 + 2 * kBfsCapMax;
constexpr int kLongRow = 192;               // rows longer than this use wave search
constexpr int kMaxLong = 256;
constexpr int kMaxBuckets = 2048;           // staged path: id buckets (in cursor[])
constexpr int kMaxBucketShift = 17;         // a bucket must fit one bitmap window
constexpr int kPiece = 256;                 // staged path: long-row part per wave step
constexpr int kMaxBound = 4;                // radix path: rows longer than cap merged, not sorted

struct BfsLds {
    alignas(16) int32_t f1[kBfsCapMax];       // capped 1-hop list (ascending)
    alignas(16) int32_t win[kWin];           // sort window, or bitmap of one id window
    int32_t cursor[kMaxRows];     // per-row position, (staged path) bucket ends, or
                                  // (radix path) the second buffer and the counters
    int32_t lrow[kMaxLong];       // rows searched/copied by whole wavefronts
    int32_t lpre[kMaxLong + 1];   // staged path: long rows' first piece; bits 16-17 the row kind
    int32_t lbeg[kMaxLong];       // staged path: long rows' first entry
    int32_t bnd_b[kMaxBound];     // radix path, level 2: rows longer than cap (first entry,
    int32_t bnd_n[kMaxBound];     // length, kind: 0 forward, 1 backward)
    int32_t bnd_k[kMaxBound];
    int32_t nbnd;
    int64_t red[kWaves];
    int64_t tot[3];               // row entries per level
    int32_t nlong;
    int32_t slot;                 // staging slot held by this workgroup, or -1
    int32_t fill;
    int32_t nout;
    int32_t n1c;
    int32_t cut_b;                // radix path: last histogram bucket below the cut
    int32_t cut_n;                // radix path: entries up to and including it
    int32_t s_lds;                // level 2: where S also is in LDS (0 none, 1 win, 2 cursor, 3 f1)
    int64_t t_start;              // wall clock at the workgroup's start
    int64_t t_level1;             // wall clock after the 1-hop level
};

// One LDS object per workgroup; helpers address it directly so the
// compiler keeps LDS (ds_*) accesses.
__shared__ BfsLds L;

// Global-memory views: graph rows and outputs are read and written through
// these, so the compiler emits global (not flat) accesses.  A flat access
// also counts against the LDS wait counter, and every LDS operation or
// shuffle that follows it waits for the memory access to complete.
typedef __attribute__((address_space(1))) const int32_t g_cint;
typedef __attribute__((address_space(1))) int32_t g_int;
#define GROW(p) ((g_cint*)(p))

__device__ __forceinline__ int lower_bound_i32(g_cint* p, int b, int e, int32_t x) {
    while (b < e) {
        int mid = (b + e) >> 1;
        if (p[mid] < x) b = mid + 1; else e = mid;
    }
    return b;
}

// First index in [b,e) with p[idx] >= x, searching forward from b.
__device__ __forceinline__ int gallop_lower_bound(g_cint* p, int b, int e, int32_t x) {
    if (b >= e || p[b] >= x) return b;
    int step = 1, lo = b;  // p[lo] < x
    while (true) {
        int nx = lo + step;
        if (nx >= e) return lower_bound_i32(p, lo + 1, e, x);
        if (p[nx] >= x) return lower_bound_i32(p, lo + 1, nx, x);
        lo = nx;
        step <<= 1;
    }
}

// First index in [b,e) with p[idx] >= x, found by the 64 lanes of a
// wavefront together: each round samples 64 split points with one load per
// lane and keeps the sub-range holding the answer (a 100k-entry row takes
// three rounds of loads instead of 17 dependent ones).  All lanes call it
// with the same arguments and get the same result.
__device__ __forceinline__ int wave_lower_bound(g_cint* p, int b, int e, int32_t x) {
    // (the count of samples below x is summed with shuffles: a ballot here
    // trips a register-class error in the ROCm 7.2 backend)
    const int lane = threadIdx.x & 63;
    while (e - b > 64) {
        const int64_t len = e - b;
        const int pos = b + (int)((len * (lane + 1)) / 65);
        int k = (p[pos] < x) ? 1 : 0;  // samples below x form a prefix
        for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
        const int nb = k == 0 ? b : b + (int)((len * k) / 65) + 1;
        const int ne = k == 64 ? e : b + (int)((len * (k + 1)) / 65) + 1;
        b = nb;
        e = ne;
    }
    int k = ((b + lane < e) && p[b + lane] < x) ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
    return b + k;
}

__device__ __forceinline__ bool row_contains(g_cint* p, int b, int e, int32_t x) {
    int k = lower_bound_i32(p, b, e, x);
    return k < e && p[k] == x;
}

__device__ int64_t block_sum(int64_t v) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if (lane == 0) L.red[wid] = v;
    __syncthreads();
    int64_t t = 0;
    for (int k = 0; k < kWaves; ++k) t += L.red[k];
    return t;
}

__device__ int32_t block_min(int32_t v) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off, 64));
    __syncthreads();
    if (lane == 0) L.red[wid] = v;
    __syncthreads();
    int64_t t = INT_MAX;
    for (int k = 0; k < kWaves; ++k) t = min(t, L.red[k]);
    return (int32_t)t;
}

// Block-wide exclusive scan of one int64 per thread; *total gets the sum.
__device__ int64_t block_excl_scan(int64_t v, int64_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) L.red[wid] = x;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int k = 0; k < kWaves; ++k) { if (k < wid) pre += L.red[k]; tot += L.red[k]; }
    *total = tot;
    return pre + x - v;
}

// The same for one int32 per thread.
__device__ int block_excl_scan32(int v, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) L.red[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        const int t = (int)L.red[k];
        if (k < wid) pre += t;
        tot += t;
    }
    *total = tot;
    return pre + x - v;
}

// block_excl_scan32 without the barrier that protects L.red from the
// previous scan's readers: for callers with a barrier since that scan.
__device__ __forceinline__ int block_excl_scan32_nb(int v, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) L.red[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        const int t = (int)L.red[k];
        if (k < wid) pre += t;
        tot += t;
    }
    *total = tot;
    return pre + x - v;
}

// Wavefront-cooperative LDS counter add: every lane with key >= 0 adds 1 to
// ctr[key]; a run of consecutive lanes with one key issues one atomic for
// the run.  All 64 lanes call it.  With RET, each such lane gets the
// counter's old value plus its rank in the run (the slot it reserved).
template <bool RET>
__device__ __forceinline__ int run_add(int32_t* ctr, int key) {
    const int lane = threadIdx.x & 63;
    const int up = __shfl_up(key, 1, 64);
    const uint64_t bnd = __ballot(lane == 0 || up != key);  // run starts
    const uint64_t upto = bnd & ((2ull << lane) - 1);
    const uint64_t above = bnd & ~((2ull << lane) - 1);
    const int start = 63 - __builtin_clzll(upto);
    const int end = above ? __builtin_ctzll(above) : 64;
    int base = 0;
    if (key >= 0 && start == lane) base = atomicAdd(&ctr[key], end - lane);
    if (!RET) return 0;
    return __shfl(base, start, 64) + (lane - start);
}

struct Query {
    CsrView g;
    int32_t root;
    int level;          // 1: rows of the root; 2: F1c list + rows of F1c
    int n1c;            // |F1c| (level 2)
    bool check_f1;      // level 2: F1 was capped, so filter F1 \ F1c
    int32_t f1max;
    BfsStage st;
    g_int* S;           // level-2 output (level 1 writes L.f1)
    bool out_lds;       // output into L.f1 (level 1; staged level 2, copied to S)
};

__device__ __forceinline__ int num_rows(const Query& q) { return q.level == 1 ? 2 : 1 + 2 * q.n1c; }

// Level 2's row 0 is the F1c list in LDS; every other row is a CSR row.
__device__ __forceinline__ bool is_f1_row(const Query& q, int r) { return q.level == 2 && r == 0; }

// CSR row r (not the F1c row): entries p[b,e) of the forward or backward
// adjacency.
__device__ __forceinline__ void graph_row(const Query& q, int r, g_cint*& p, int& b, int& e) {
    int32_t v;
    int d;
    if (q.level == 1) {
        v = q.root; d = r;
    } else {
        v = L.f1[(r - 1) >> 1]; d = (r - 1) & 1;
    }
    g_cint* off = GROW(d ? q.g.bwd_off : q.g.fwd_off);
    p = GROW(d ? q.g.bwd_adj : q.g.fwd_adj);
    b = off[v];
    e = off[v + 1];
}

// Row r's bounds (the F1c row: [0, n1c) of L.f1).
__device__ __forceinline__ void row_bounds(const Query& q, int r, g_cint*& p, int& b, int& e) {
    if (is_f1_row(q, r)) { p = nullptr; b = 0; e = q.n1c; return; }
    graph_row(q, r, p, b, e);
}

__device__ __forceinline__ void put_out(const Query& q, int pos, int32_t x) {
    if (q.out_lds) L.f1[pos] = x; else q.S[pos] = x;
}

__device__ __forceinline__ bool keep_id(const Query& q, int32_t x) {
    if (x == q.root) return false;
    if (q.level == 2 && q.check_f1 && x > q.f1max) {
        g_cint* fo = GROW(q.g.fwd_off); g_cint* bo = GROW(q.g.bwd_off);
        if (row_contains(GROW(q.g.fwd_adj), fo[q.root], fo[q.root + 1], x)) return false;
        if (row_contains(GROW(q.g.bwd_adj), bo[q.root], bo[q.root + 1], x)) return false;
    }
    return true;
}

// Appends the set bits of bm[0,nw) (ids lo + bit index) in ascending order
// to the output after the L.nout already there, up to cap; updates L.nout.
// Each thread takes a contiguous run of words.  All threads call it.
// Each thread holds kXW consecutive words in registers (four 16-byte LDS
// loads), so kWin / kXW threads cover a full window.
constexpr int kXW = 16;
static_assert(kWin % kXW == 0 && kWin / kXW <= kBT, "bitmap window per thread");
static_assert(kRadixMax <= kWin && kRadixMax + 2 * kBT <= kMaxRows && kRadixMax % kBT == 0 &&
                  kHistBuckets <= kRadixMax,
              "radix path buffers");
__device__ __forceinline__ void extract_bits(const Query& q, const uint32_t* bm, int nw, int32_t lo,
                                             int cap) {
    const int w0 = threadIdx.x * kXW;
    uint32_t wv[kXW];
    if (w0 < nw) {
#pragma unroll
        for (int k = 0; k < kXW / 4; ++k) {
            const uint4 v = reinterpret_cast<const uint4*>(bm + w0)[k];
            wv[4 * k] = v.x; wv[4 * k + 1] = v.y; wv[4 * k + 2] = v.z; wv[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < kXW; ++k)
            if (w0 + k >= nw) wv[k] = 0u;  // past the window: stale words
    } else {
#pragma unroll
        for (int k = 0; k < kXW; ++k) wv[k] = 0u;
    }
    int cntw = 0;
#pragma unroll
    for (int k = 0; k < kXW; ++k) cntw += __popc(wv[k]);
    int tot;
    int pos = block_excl_scan32(cntw, &tot) + L.nout;
    if (cntw > 0 && pos < cap) {
#pragma unroll
        for (int k = 0; k < kXW; ++k) {
            uint32_t bits = wv[k];
            while (bits && pos < cap) {
                const int bit = __builtin_ctz(bits);
                bits &= bits - 1;
                put_out(q, pos++, lo + ((w0 + k) << 5) + bit);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int no = L.nout + tot;
        L.nout = no < cap ? no : cap;
    }
    __syncthreads();
}

// Visits entries of CSR rows 64 rows per wavefront: lane k holds row r0+k
// (its first entry and length, lengths cut by the caller), the chunk's
// positions are spread lane-strided, each lane finding its row by a binary
// search over the lanes' starts, so consecutive lanes read consecutive
// entries with kG loads in flight per lane.  visit(x, pos, ok) gets the entry
// and its position in the chunk, and is called by all lanes together with
// ok = whether the lane holds an entry; returns the chunk's total.
template <typename Visit>
__device__ __forceinline__ int chunk_rows(g_cint* pbase, int b, int len, Visit visit) {
    const int lane = threadIdx.x & 63;
    int x = len;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int pos = x - len;
    const int chunk = __shfl(x, 63, 64);
    const uint64_t pa = (uint64_t)(pbase + b);
    constexpr int kG = 8;
    for (int p0 = 0; p0 < chunk; p0 += 64 * kG) {
        int32_t val[kG];
#pragma unroll
        for (int g2 = 0; g2 < kG; ++g2) {
            val[g2] = 0;
            if (p0 + g2 * 64 >= chunk) continue;  // uniform: past the chunk
            const int pp = p0 + g2 * 64 + lane;
            int rr = 0;
#pragma unroll
            for (int st = 32; st > 0; st >>= 1) {
                const int sp = __shfl(pos, rr + st, 64);
                if (sp <= pp) rr += st;
            }
            const int k = pp - __shfl(pos, rr, 64);
            const uint64_t src = __shfl(pa, rr, 64);
            if (pp < chunk) val[g2] = reinterpret_cast<g_cint*>(src)[k];
        }
#pragma unroll
        for (int g2 = 0; g2 < kG; ++g2) {
            if (p0 + g2 * 64 >= chunk) continue;
            const int pp = p0 + g2 * 64 + lane;
            visit(val[g2], pp, pp < chunk);
        }
    }
    return chunk;
}

// Staged windows: as extract_bits, with the count kept by the caller (nout,
// the same in every thread; returns the new count) and the words read
// zeroed for the next window.  One barrier: the caller has one between the
// window's bitmap writes and this call, which also separates it from the
// previous call's reads of L.red.
__device__ __forceinline__ int extract_bits_staged(const Query& q, uint32_t* bm, int nw, int32_t lo,
                                                   int cap, int nout) {
    const int w0 = threadIdx.x * kXW;
    uint32_t wv[kXW];
    if (w0 < nw) {
#pragma unroll
        for (int k = 0; k < kXW / 4; ++k) {
            const uint4 v = reinterpret_cast<const uint4*>(bm + w0)[k];
            wv[4 * k] = v.x; wv[4 * k + 1] = v.y; wv[4 * k + 2] = v.z; wv[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < kXW / 4; ++k) reinterpret_cast<uint4*>(bm + w0)[k] = make_uint4(0, 0, 0, 0);
    } else {
#pragma unroll
        for (int k = 0; k < kXW; ++k) wv[k] = 0u;
    }
    int cntw = 0;
#pragma unroll
    for (int k = 0; k < kXW; ++k) cntw += __popc(wv[k]);
    int tot;
    int pos = block_excl_scan32_nb(cntw, &tot) + nout;
    if (cntw > 0 && pos < cap) {
#pragma unroll
        for (int k = 0; k < kXW; ++k) {
            uint32_t bits = wv[k];
            while (bits && pos < cap) {
                const int bit = __builtin_ctz(bits);
                bits &= bits - 1;
                put_out(q, pos++, lo + ((w0 + k) << 5) + bit);
            }
        }
    }
    const int no = nout + tot;
    return no < cap ? no : cap;
}

// Staged path: visits every entry of the level's rows, each row cut to its
// first cap+1 entries (a longer row's entry [cap] already bounds the
// answer): the F1c list from LDS, CSR rows up to kLongRow entries through
// chunk_rows, longer ones (L.lrow, pieces in L.lpre/L.lbeg) kPiece entries
// per wavefront step.  All threads call it; visit(x, ok) is called by whole
// wavefronts, ok telling the lanes that hold an entry.
template <typename Visit>
__device__ __forceinline__ void stage_rows(const Query& q, int nr, int cap, bool long_ok,
                                           int nlong, int npieces, Visit visit,
                                           int skip_over = INT_MAX) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (q.level == 2)
        for (int k0 = 0; k0 < q.n1c; k0 += kBT) {
            const int k = k0 + (int)threadIdx.x;
            visit(k < q.n1c ? L.f1[k] : 0, k < q.n1c);
        }
    for (int r0 = wid * 64; r0 < nr; r0 += kBT) {
        const int r = r0 + lane;
        g_cint* p = GROW(q.g.fwd_adj);
        int b = 0, e = 0;
        if (r < nr && !is_f1_row(q, r)) graph_row(q, r, p, b, e);
        int len = min(e - b, cap + 1);
        if (long_ok && e - b > kLongRow) len = 0;
        if (e - b > skip_over) len = 0;  // merged afterwards (merge_bound)
        chunk_rows(p, b, len, [&](int32_t x, int, bool ok) { visit(x, ok); });
    }
    // kPP pieces per wavefront step: kPP * kPiece / 64 loads in flight per lane
    constexpr int kPP = 4;
    constexpr int kPL = kPiece / 64;
    for (int pc0 = wid; pc0 < npieces; pc0 += kPP * kWaves) {
        int32_t val[kPP][kPL];
        int lenr[kPP], off[kPP];
#pragma unroll
        for (int j = 0; j < kPP; ++j) {
            const int pc = pc0 + j * kWaves;
            lenr[j] = 0;
            off[j] = 0;
            if (pc < npieces) {
                int a = 0, z = nlong;  // long row with first piece <= pc < next row's
                while (z - a > 1) {
                    const int mid = (a + z) >> 1;
                    if ((L.lpre[mid] & 0xffff) <= pc) a = mid; else z = mid;
                }
                const int pj = L.lpre[a];
                g_cint* p = GROW(((pj >> 16) & 1) ? q.g.bwd_adj : q.g.fwd_adj) + L.lbeg[a];
                lenr[j] = pj >> 18;
                off[j] = (pc - (pj & 0xffff)) * kPiece;
#pragma unroll
                for (int g2 = 0; g2 < kPL; ++g2) {
                    const int k = off[j] + g2 * 64 + lane;
                    val[j][g2] = k < lenr[j] ? p[k] : 0;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kPP; ++j)
#pragma unroll
            for (int g2 = 0; g2 < kPL; ++g2)
                visit(val[j][g2], off[j] + g2 * 64 + lane < lenr[j]);
    }
}

// Value of lane (lane ^ J) without the LDS crossbar: DPP quad permutes and
// row mirrors for J < 16 (xor 4 = half-row mirror then xor 3; xor 8 = row
// mirror then half-row mirror) and gfx950's row-pair swaps for 16 and 32.
template <int J>
__device__ __forceinline__ int32_t xor_lane(int32_t v) {
    static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "lane xor");
    if constexpr (J == 1) {
        return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    } else if constexpr (J == 4) {
        const int32_t t = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);  // ^7
        return __builtin_amdgcn_mov_dpp(t, 0x1B, 0xF, 0xF, true);             // ^3
    } else if constexpr (J == 8) {
        const int32_t t = __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true);  // ^15
        return __builtin_amdgcn_mov_dpp(t, 0x141, 0xF, 0xF, true);            // ^7
    } else if constexpr (J == 16) {
        // {v with odd rows <- even rows, v with even rows <- odd rows}
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return ((threadIdx.x >> 4) & 1) ? (int32_t)r[0] : (int32_t)r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return ((threadIdx.x >> 5) & 1) ? (int32_t)r[0] : (int32_t)r[1];
    }
}

// Median of three, written so that the compiler emits one v_med3_i32: with
// c = INT_MIN it is min(a, b), with c = INT_MAX max(a, b).
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
    return min(max(a, b), max(min(a, b), c));
}

// Hides x's value (a per-lane INT_MIN/INT_MAX bound) from the compiler, so
// that a v_med3 against it stays one instruction instead of a min, a max
// and a select.
__device__ __forceinline__ void opaque(int32_t& x) { asm("" : "+v"(x)); }


// Bitonic sort of s[0,P) (P = 64E a power of two <= 4096) by one wavefront,
// in registers, lane-major: lane l holds elements [lE, lE+E) in registers
// 0..E-1.  A compare distance J below E pairs two registers of one lane (a
// min and a max, or two v_med3 where the direction depends on the lane); a
// distance J >= E pairs register e of lanes l and l ^ J/E (xor_lane: DPP and
// row swaps) with one v_med3 against a per-lane bound.  Every merge ends
// with its log2(E) in-lane steps, so for E >= 8 most of the network needs no
// lane exchange.  Loads and stores are 16-byte pieces (E >= 4); they touch
// other lanes' natural-order elements, hence the fences on both sides.
// One compare-exchange step (distance J) of merge size K, recursing down to
// J = 1; compile-time J and K, so every lane exchange is a fixed pattern.
template <int E, int K, int J>
__device__ __forceinline__ void bitonic_step(int32_t (&v)[E], int lane) {
    if constexpr (J < E) {
        // element lE+e ascends iff bit K of it is clear: of e for K < E,
        // of the lane (bit K/E) otherwise
        const bool up_l = (lane & (K >= E ? K / E : 0)) == 0;
        int32_t b_lo = up_l ? INT_MIN : INT_MAX, b_hi = up_l ? INT_MAX : INT_MIN;
        opaque(b_lo);
        opaque(b_hi);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if ((e & J) == 0) {
                const int32_t x = v[e], y = v[e + J];
                if constexpr (K < E) {
                    const bool up = (e & K) == 0;
                    v[e] = up ? min(x, y) : max(x, y);
                    v[e + J] = up ? max(x, y) : min(x, y);
                } else {
                    v[e] = med3_i32(x, y, b_lo);
                    v[e + J] = med3_i32(x, y, b_hi);
                }
            }
        }
    } else {
        constexpr int d = J / E;
        // keep the minimum where "lower of the pair" equals "ascending"
        const bool lower = (lane & d) == 0;
        const bool up = (lane & (K / E)) == 0;
        int32_t b = (lower == up) ? INT_MIN : INT_MAX;
        opaque(b);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = med3_i32(v[e], xor_lane<d>(v[e]), b);
    }
    if constexpr (J > 1) bitonic_step<E, K, J / 2>(v, lane);
}

template <int E, int K>
__device__ __forceinline__ void bitonic_merges(int32_t (&v)[E], int lane) {
    bitonic_step<E, K, K / 2>(v, lane);
    if constexpr (K < 64 * E) bitonic_merges<E, 2 * K>(v, lane);
}

// OPQ: the lane index is opaque, so the per-lane bounds are computed where
// they are used; otherwise the compiler hoists them to the kernel's start,
// which in the workgroup kernel (128 VGPRs, two call sites) spills them.
template <int E, bool OPQ = false>
__device__ __forceinline__ void wave_sort_reg(int32_t* s) {
    int lane = (int)threadIdx.x & 63;
    if constexpr (OPQ) asm volatile("" : "+v"(lane));
    int32_t v[E];
    int32_t* mine = s + lane * E;
    wave_lds_sync();
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int e = 0; e < E; e += 4) {
            const int4 q = *reinterpret_cast<const int4*>(mine + e);
            v[e] = q.x; v[e + 1] = q.y; v[e + 2] = q.z; v[e + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mine[e];
    }
    bitonic_merges<E, 2>(v, lane);
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int e = 0; e < E; e += 4)
            *reinterpret_cast<int4*>(mine + e) = make_int4(v[e], v[e + 1], v[e + 2], v[e + 3]);
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) mine[e] = v[e];
    }
    wave_lds_sync();
}

// One wavefront: the distinct ids of the sorted s[0,n) other than root, the
// first limit of them into out (another array); returns their number,
// capped.
__device__ __forceinline__ int wave_unique_to(const int32_t* s, int n, int limit, int32_t root,
                                              int32_t* out) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int cnt = 0;
    int32_t carry = INT_MIN;  // last element of the previous chunk
    for (int base = 0; base < n && cnt < limit; base += 64) {
        const int i = base + lane;
        const int32_t x = i < n ? s[i] : INT_MAX;
        int32_t prev = __shfl_up(x, 1, 64);
        if (lane == 0) prev = carry;
        const bool k = i < n && x != prev && x != root;
        const uint64_t mask = __ballot(k);
        const int pos = cnt + __popcll(mask & below);
        carry = __shfl(x, 63, 64);
        if (k && pos < limit) out[pos] = x;
        cnt += __popcll(mask);
    }
    return cnt < limit ? cnt : limit;
}

// Radix path, step 1: the level's row entries (rows cut as in stage_rows)
// below lim, the root dropped, appended to A[0, kRadixMax) in visit order;
// returns their number (those past the capacity are counted, not stored).
// With H (kHistBuckets zeroed counters), also counts them per bucket
// x >> hsh.  All threads call it.
__device__ __forceinline__ int radix_gather(const Query& q, int nr, int cap, bool long_ok, int nlong,
                                            int npieces, int32_t lim, int32_t* A, int32_t* H,
                                            int hsh, int skip_over) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    if (threadIdx.x == 0) L.fill = 0;
    __syncthreads();
    stage_rows(q, nr, cap, long_ok, nlong, npieces, [&](int32_t x, bool ok) {
        const bool keep = ok && x < lim && x != q.root;
        const uint64_t m = __ballot(keep);
        if (m == 0) return;  // uniform over the wavefront
        int base = 0;
        if (lane == 0) base = atomicAdd(&L.fill, __popcll(m));
        base = __shfl(base, 0, 64);
        const int at = base + __popcll(m & lt);
        if (keep && at < kRadixMax) A[at] = x;
        if (H) run_add<false>(H, keep ? x >> hsh : -1);
    }, skip_over);
    __syncthreads();
    return L.fill;
}

// Radix path, step 2: stable LSD radix sort of A[0,M) (M <= kRadixMax),
// 8-bit digits over the low `bits` bits, B[0,M) the other buffer; returns
// the buffer with the result.  Wavefront w holds entries [w*64K, (w+1)*64K)
// (K = ceil(M / kBT)), round r at lane l entry w*64K + 64r + l; an entry's
// rank among its wavefront's entries of the same digit comes from eight
// ballots (the lanes with equal digits) and a per-wavefront counter, the
// counters are scanned in (digit, wavefront) order, and each entry is
// stored at its digit's offset plus its rank.  The counters are 16-bit, two
// per word of C[0, kBT*2), stored wavefront-major (wavefront w's counter
// for digit d at half d & 1 of word 128w + d/2): a wavefront's lanes then
// spread over 128 consecutive words, where digit-major storage (word
// 4d + w/2) put all of them on one LDS bank in four.  All threads call it.
__device__ __forceinline__ int32_t* radix_sort(int32_t* A, int32_t* B, uint32_t* C, int M,
                                               int bits) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int K = (M + kBT - 1) / kBT;
    const int w0 = wid * 64 * K;
    const uint64_t lt = (1ull << lane) - 1;
    for (int shift = 0; shift < bits; shift += 8) {
        C[threadIdx.x] = 0u;
        C[threadIdx.x + kBT] = 0u;
        __syncthreads();
        int rk[kRadixRounds];  // the keys are read again for the scatter
#pragma unroll
        for (int r = 0; r < kRadixRounds; ++r) {
            rk[r] = 0;
            if (r < K) {  // uniform
                const int idx = w0 + r * 64 + lane;
                const bool valid = idx < M;
                const uint32_t d = valid ? ((uint32_t)A[idx] >> shift) & 255u : 0u;
                uint64_t peers = __ballot(valid);
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const bool bit = (d >> b) & 1u;
                    const uint64_t mb = __ballot(bit);
                    peers &= bit ? mb : ~mb;
                }
                const int wi = wid * 128 + (int)(d >> 1);
                const int hs = (int)(d & 1u) * 16;
                const int before = (int)((C[wi] >> hs) & 0xFFFFu);
                rk[r] = before + __popcll(peers & lt);
                if (valid && (peers & lt) == 0)
                    atomicAdd(&C[wi], (uint32_t)__popcll(peers) << hs);
            }
        }
        __syncthreads();
        {
            // thread t scans digit t/2's counters of wavefronts 4(t&1) to
            // 4(t&1)+3 (all in one half of their words); the thread index
            // is opaque so the four addresses are not hoisted and spilled
            int tid = (int)threadIdx.x;
            asm volatile("" : "+v"(tid));
            const int dg = tid >> 1, w0c = (tid & 1) * 4;
            const int hs = (dg & 1) * 16;
            int c[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = (int)((C[(w0c + k) * 128 + (dg >> 1)] >> hs) & 0xFFFFu);
            int tot;
            int pk = block_excl_scan32(c[0] + c[1] + c[2] + c[3], &tot);
            // a word's two digits belong to threads t and t ^ 2 (one
            // wavefront, which alone reads the word): the even digit's
            // thread writes both offsets
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int other = __shfl_xor(pk, 2, 64);
                if (!hs) C[(w0c + k) * 128 + (dg >> 1)] = (uint32_t)pk | ((uint32_t)other << 16);
                pk += c[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRadixRounds; ++r) {
            const int idx = w0 + r * 64 + lane;
            if (r < K && idx < M) {
                const int32_t key = A[idx];
                const uint32_t d = ((uint32_t)key >> shift) & 255u;
                B[(int)((C[wid * 128 + (int)(d >> 1)] >> ((d & 1u) * 16)) & 0xFFFFu) + rk[r]] = key;
            }
        }
        __syncthreads();
        int32_t* t = A;
        A = B;
        B = t;
    }
    return A;
}

// Radix path, step 3: the distinct kept ids of the sorted Z[0,M), the
// first cap of them to the output (or to dst in LDS, another array);
// returns how many there are (uncapped).  Each thread takes a contiguous
// run.  All threads call it.
__device__ __forceinline__ int radix_emit(const Query& q, const int32_t* Z, int M, int cap,
                                          int32_t* dst = nullptr, int32_t* also = nullptr) {
    const int chunk = (M + kBT - 1) / kBT;
    const int s0 = min(M, (int)threadIdx.x * chunk);
    const int s1 = min(M, s0 + chunk);
    int mine = 0;
    for (int k = s0; k < s1; ++k) {
        const int32_t x = Z[k];
        if ((k == 0 || Z[k - 1] != x) && keep_id(q, x)) ++mine;
    }
    int tot;
    int pos = block_excl_scan32(mine, &tot);
    for (int k = s0; k < s1 && pos < cap; ++k) {
        const int32_t x = Z[k];
        if ((k == 0 || Z[k - 1] != x) && keep_id(q, x)) {
            if (dst) dst[pos] = x; else put_out(q, pos, x);
            if (also) also[pos] = x;
            ++pos;
        }
    }
    __syncthreads();  // the output is complete
    return tot;
}

// One step of merge_bound: Y[0, min(result, cap)) = the distinct kept ids
// of the ascending X[0,np) and F[0,nf) together; returns their number
// (uncapped).  Each thread takes a contiguous run of the merged order,
// found by a merge-path search (X first among equal entries), walks it
// once to count and once to write.  All threads call it.
__device__ __forceinline__ int merge_step(const Query& q, const int32_t* X, int np,
                                          const int32_t* F, int nf, int cap, int32_t* Y) {
    const int T = np + nf;
    const int per = (T + kBT - 1) / kBT;
    const int d0 = min(T, (int)threadIdx.x * per), d1 = min(T, d0 + per);
    int a = max(0, d0 - nf), z = min(d0, np);
    while (a < z) {
        const int mid = (a + z) >> 1;
        if (X[mid] <= F[d0 - 1 - mid]) a = mid + 1; else z = mid;
    }
    const int i0 = a, j0 = d0 - a;
    // the merged entry before the run: the larger of the last ones taken
    const int32_t p0 = max(i0 > 0 ? X[i0 - 1] : INT_MIN, j0 > 0 ? F[j0 - 1] : INT_MIN);
    int i = i0, j = j0, mine = 0;
    int32_t prev = p0;
    for (int d = d0; d < d1; ++d) {
        const bool tx = j >= nf || (i < np && X[i] <= F[j]);
        const int32_t x = tx ? X[i] : F[j];
        i += tx; j += !tx;
        if (x != prev && keep_id(q, x)) ++mine;
        prev = x;
    }
    int tot;
    int pos = block_excl_scan32(mine, &tot);
    i = i0; j = j0; prev = p0;
    for (int d = d0; d < d1 && pos < cap; ++d) {
        const bool tx = j >= nf || (i < np && X[i] <= F[j]);
        const int32_t x = tx ? X[i] : F[j];
        i += tx; j += !tx;
        if (x != prev && keep_id(q, x)) Y[pos++] = x;
        prev = x;
    }
    __syncthreads();  // Y is complete
    return tot;
}

// Radix path with the long rows (L.bnd_*, nbnd of them: at level 2 those
// longer than cap, at level 1 those above kWaveSortMax) left out of the
// gather: the distinct kept ids of the sorted other entries Z[0,M) (in
// L.win or L.cursor), then each long row's entries below the bound merged
// in; the first cap to the output; returns their number (capped).  A row
// is ascending in memory, so it needs no sort, and only its entries below
// the bound among its first cap+1 are read.  The running
// list alternates between L.win[0,cap) and L.cursor[0,cap); a row's entries
// are staged at L.cursor[kBfsCapMax, kBfsCapMax + cap + 1).  All threads
// call it.
static_assert(2 * kBfsCapMax + 1 <= kMaxRows && kBfsCapMax <= kWin, "merge_bound buffers");
__device__ __forceinline__ int merge_bound(const Query& q, const int32_t* Z, int M, int cap,
                                           int32_t lim, int nbnd) {
    int32_t* X = (Z == L.win) ? L.cursor : L.win;
    int32_t* Y = (Z == L.win) ? L.win : L.cursor;
    int32_t* F = L.cursor + kBfsCapMax;
    // a row's entries are loaded into registers one row ahead, so the loads
    // are in flight during the previous row's merge
    constexpr int kFP = (kBfsCapMax + kBT) / kBT;  // cap + 1 entries per row at most
    int32_t fv[kFP];
    auto fetch = [&](int k) {
        g_cint* p = GROW(L.bnd_k[k] ? q.g.bwd_adj : q.g.fwd_adj) + L.bnd_b[k];
        const int len = L.bnd_n[k];
#pragma unroll
        for (int u = 0; u < kFP; ++u) {
            const int t = (int)threadIdx.x + u * kBT;
            fv[u] = t < len ? p[t] : 0;
        }
    };
    fetch(0);
    int np = min(radix_emit(q, Z, M, cap, X), cap);
    for (int k = 0; k < nbnd; ++k) {
        // ids above the cap-th of the running list cannot enter it
        const int32_t hi = np == cap ? min(lim, X[cap - 1]) : lim;
        const int len = L.bnd_n[k];
#pragma unroll
        for (int u = 0; u < kFP; ++u) {
            const int t = (int)threadIdx.x + u * kBT;
            if (t < len) F[t] = fv[u];
        }
        __syncthreads();
        if (k + 1 < nbnd) fetch(k + 1);
        int a = 0, z = len;  // entries below hi
        while (a < z) {
            const int mid = (a + z) >> 1;
            if (F[mid] < hi) a = mid + 1; else z = mid;
        }
        np = min(merge_step(q, X, np, F, a, cap, Y), cap);
        int32_t* t = X;
        X = Y;
        Y = t;
    }
    for (int t = threadIdx.x; t < np; t += kBT) put_out(q, t, X[t]);
    if (threadIdx.x == 0) L.s_lds = X == L.win ? 1 : 2;
    __syncthreads();  // the output is complete
    return np;
}

// Collects the `cap` smallest distinct kept ids of the level's rows into the
// output (level 1: L.f1, level 2: q.S; ascending); returns the count.  All
// threads call it.  Row entries below the answer bound (from rows longer
// than cap) are gathered into LDS and radix-sorted there: always up to
// kSortMax row entries, and (q.st.radix) any list whose entries below the
// bound, or below a tighter bound from their histogram, fit kRadixMax.
// Other lists are staged in a pool slot bucketed by id when they fit, else
// swept through the id range in bitmap windows; either way ids come out of
// LDS bitmaps in ascending order.  It has one call site (two_hop) in one
// kernel (k_bfs_root): out-of-line calls cost a scratch stack frame per lane.
__device__ __forceinline__ int collect(Query& q, int cap) {
    const int32_t n = (int32_t)q.g.n;
    const int nr = num_rows(q);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (q.level == 1) {
        // a 1-hop list of at most kWaveSortMax entries: the first wavefront
        // loads both rows, sorts them in registers and compacts them into
        // L.f1; the others wait at one barrier
        g_cint* fo = GROW(q.g.fwd_off);
        g_cint* bo = GROW(q.g.bwd_off);
        const int32_t fb = fo[q.root], fe = fo[q.root + 1], bb = bo[q.root], be = bo[q.root + 1];
        const int df = fe - fb, d1 = df + (be - bb);
        if (d1 <= kWaveSortMax) {
            if (wid == 0) {
                int P = 64;
                while (P < d1) P <<= 1;
                g_cint* fa = GROW(q.g.fwd_adj);
                g_cint* ba = GROW(q.g.bwd_adj);
                for (int k = lane; k < P; k += 64)
                    L.win[k] = k < df ? fa[fb + k] : (k < d1 ? ba[bb + k - df] : INT_MAX);
                switch (P) {
                    case 64: wave_sort_reg<1, true>(L.win); break;
                    case 128: wave_sort_reg<2, true>(L.win); break;
                    case 256: wave_sort_reg<4, true>(L.win); break;
                    default: wave_sort_reg<8, true>(L.win); break;
                }
                const int n1 = wave_unique_to(L.win, d1, cap, q.root, L.f1);
                if (lane == 0) {
                    L.nout = n1;
                    L.tot[1] = d1;
                }
            }
            __syncthreads();
            return L.nout;
        }
    }
    if (threadIdx.x == 0) { L.nlong = 0; L.nbnd = 0; L.s_lds = 0; }
    __syncthreads();
    // row entries (bits 26 and up) and entries cut at cap+1 per row (bits
    // 0-25: at most kMaxRows * (kBfsCapMax + 1) < 2^26)
    int64_t mine = 0;
    int32_t tb = INT_MAX;
    // rows longer than mthr are merged (merge_bound) when there are at most
    // kMaxBound of them: at level 2 those longer than cap (they bound the
    // answer); at level 1 (two rows, more than kWaveSortMax entries
    // together) any row above kWaveSortMax
    const int mthr = q.level == 1 ? kWaveSortMax : cap;
    for (int r = threadIdx.x; r < nr; r += kBT) {
        g_cint* p; int b, e;
        row_bounds(q, r, p, b, e);
        L.cursor[r] = b;
        mine += ((int64_t)(e - b) << 26) + min(e - b, cap + 1);
        if (is_f1_row(q, r)) continue;  // at most cap entries, in LDS
        // a row with more than cap entries holds at least cap kept ids (only
        // the root can be dropped from it) among its first cap+1, so no id
        // above its entry [cap] can be in the answer
        if (e - b > cap) tb = min(tb, p[b + cap]);
        if (e - b > mthr) {
            const int s = atomicAdd(&L.nbnd, 1);
            if (s < kMaxBound) {
                L.bnd_b[s] = b;
                L.bnd_n[s] = min(e - b, cap + 1);
                L.bnd_k[s] = q.level == 1 ? r : (r - 1) & 1;
            }
        }
        if (e - b > kLongRow) {
            const int s = atomicAdd(&L.nlong, 1);
            if (s < kMaxLong) L.lrow[s] = r;
        }
    }
    const int64_t packed = block_sum(mine);
    const int64_t total = packed >> 26;
    const int64_t total_cut = packed & ((int64_t(1) << 26) - 1);
    tb = block_min(tb);
    // with a capped 1-hop list, F1c alone holds cap kept ids, all <= f1max;
    // below the bound the F1 \ F1c filter never applies (those ids exceed
    // f1max), so only the root is cleared from the bitmap
    if (q.level == 2 && q.n1c == cap && q.n1c > 0) tb = min(tb, q.f1max);
    const int64_t span = (tb == INT_MAX) ? (int64_t)n : (int64_t)tb + 1;
    int sh = 0;
    while (((span - 1) >> sh) >= kMaxBuckets) ++sh;
    const bool want_stage = total > kSortMax && total <= q.st.max_total &&
                            total_cut <= q.st.cap_entries && sh <= kMaxBucketShift &&
                            q.st.n_slots > 0;
    if (threadIdx.x == 0) {
        L.nout = 0; L.tot[q.level] = total;
        if (L.nlong > kMaxLong) L.nlong = -1;  // list overflow: every row per thread
        L.cut_b = -1;
    }
    __syncthreads();
    if (total == 0) return 0;
    const bool long_ok = L.nlong >= 0;
    const int nlong = long_ok ? L.nlong : 0;
    const bool radix = total <= kSortMax || q.st.radix;
    // the rows longer than mthr (at most kMaxBound of them) are merged into
    // the sorted other entries, not gathered: at level 1 always, at level 2
    // on the radix path
    const int nbnd = L.nbnd;
    const bool merge = (q.level == 1 || radix) && nbnd > 0 && nbnd <= kMaxBound;
    const int skip_over = merge ? mthr : INT_MAX;
    int64_t merged_cut = 0;  // their entries, cut at cap+1
    if (merge)
        for (int k = 0; k < nbnd; ++k) merged_cut += L.bnd_n[k];
    // long rows in pieces of kPiece entries: first piece, kind and length
    // per long row (L.lpre), first entry (L.lbeg); with skip, the rows
    // longer than cap get none
    auto set_pieces = [&](bool skip) {
        // (an opaque copy of the thread index: left visible, the LDS
        // addresses below are hoisted out of the level loop and spilled)
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        int cnt = 0, pk = 0, bb = 0, ln = 0;
        if (tid < nlong) {
            const int r = L.lrow[tid];
            g_cint* p; int b, e;
            graph_row(q, r, p, b, e);
            ln = (skip && e - b > mthr) ? 0 : min(e - b, cap + 1);
            cnt = (ln + kPiece - 1) / kPiece;
            pk = q.level == 1 ? r : (r - 1) & 1;
            bb = b;
        }
        int64_t tp;
        const int pre = (int)block_excl_scan(cnt, &tp);
        if (tid < nlong) {
            L.lpre[tid] = pre | (pk << 16) | (ln << 18);
            L.lbeg[tid] = bb;
        }
        return (int)tp;
    };
    int npieces = set_pieces(merge);
    if (radix || merge) {
        // radix path: the entries below the bound gathered into L.win and
        // sorted there.  When more than kRadixMax of them, a histogram of
        // the first gather gives a lower bound holding cap + cap/4 entries
        // and they are gathered again below it; kept ids short of cap
        // below such a bound (duplicates) leave it to the paths below.  In
        // merge mode only the short rows' entries are gathered and sorted
        // (merge_bound adds the long rows').
        int32_t* A = L.win;
        int32_t* B = L.cursor;
        uint32_t* C = reinterpret_cast<uint32_t*>(L.cursor + kRadixMax);
        int32_t lim = (int32_t)span;
        const bool hist = total_cut - merged_cut > kRadixMax;
        int hsh = 0;
        while (((lim - 1) >> hsh) >= kHistBuckets) ++hsh;
        if (hist)
            for (int k = threadIdx.x; k < kHistBuckets; k += kBT) B[k] = 0;
        int M = radix_gather(q, nr, cap, long_ok, nlong, npieces, lim, A, hist ? B : nullptr, hsh,
                             skip_over);
        bool cut = false;
        if (M > kRadixMax) {
            static_assert(kHistBuckets == 2 * kBT, "two buckets per thread");
            const int target = cap + cap / 4;
            const int v0 = B[2 * threadIdx.x], v1 = B[2 * threadIdx.x + 1];
            int tot;
            const int p0 = block_excl_scan32(v0 + v1, &tot);
            if (p0 < target && p0 + v0 + v1 >= target) {
                const bool first = p0 + v0 >= target;
                L.cut_b = 2 * (int)threadIdx.x + (first ? 0 : 1);
                L.cut_n = first ? p0 + v0 : p0 + v0 + v1;
            }
            __syncthreads();
            const int cb = L.cut_b;
            if (cb >= 0 && L.cut_n <= kRadixMax) {
                lim = (int32_t)min((int64_t)lim, (int64_t)(cb + 1) << hsh);
                cut = true;
                M = radix_gather(q, nr, cap, long_ok, nlong, npieces, lim, A, nullptr, 0, skip_over);
            }
        }
        if (M <= kRadixMax) {
            const int32_t* Z = A;
            if (M <= kWaveSortMax) {
                // few entries: one wavefront's register sort, no barriers
                int P = 64;
                while (P < M) P <<= 1;
                for (int k = M + (int)threadIdx.x; k < P; k += kBT) A[k] = INT_MAX;
                __syncthreads();
                if (wid == 0) {
                    switch (P) {
                        case 64: wave_sort_reg<1, true>(A); break;
                        case 128: wave_sort_reg<2, true>(A); break;
                        case 256: wave_sort_reg<4, true>(A); break;
                        default: wave_sort_reg<8, true>(A); break;
                    }
                }
                __syncthreads();
            } else {
                const int bits = 32 - __builtin_clz((uint32_t)max(lim - 1, 1));
                Z = radix_sort(A, B, C, M, bits);
            }
            if (merge) {
                const int kept = merge_bound(q, Z, M, cap, lim, nbnd);
                if (!cut || kept >= cap) return kept;
            } else {
                // (a copy of S in the buffer Z is not in, for the CDF)
                int32_t* also = q.level == 2 ? (Z == L.win ? L.cursor : L.win) : nullptr;
                if (threadIdx.x == 0 && also) L.s_lds = Z == L.win ? 2 : 1;
                const int kept = radix_emit(q, Z, M, cap, nullptr, also);
                if (!cut || kept >= cap) return min(kept, cap);
            }
        }
        if (merge) npieces = set_pieces(false);
        if (threadIdx.x == 0) L.s_lds = 0;  // the paths below write S afresh
        // the sweep's row cursors again
        for (int r = threadIdx.x; r < nr; r += kBT) {
            g_cint* p; int b, e;
            row_bounds(q, r, p, b, e);
            L.cursor[r] = b;
        }
    }
    if (threadIdx.x == 0) {
        L.nout = 0;
        if (want_stage && L.slot < 0) {
            // first free slot from a block-dependent start; one pass at most
            for (int t = 0; t < q.st.n_slots; ++t) {
                const int k = (int)((blockIdx.x + (uint32_t)t) % (uint32_t)q.st.n_slots);
                if (atomicCAS(&q.st.busy[k], 0, 1) == 0) { L.slot = k; break; }
            }
            __threadfence();
        }
    }
    __syncthreads();
    uint32_t* bm = reinterpret_cast<uint32_t*>(L.win);
    if (want_stage && L.slot >= 0) {
        // staged: the entries below the bound are counted per id bucket
        // (width 2^sh), then written bucket by bucket into the slot, and
        // the windows are runs of whole buckets, each reading its own
        // entries only
        g_int* buf = (g_int*)(q.st.buf + (int64_t)L.slot * q.st.cap_entries);
        const int nbk = (int)((span - 1) >> sh) + 1;
        int32_t* E = L.cursor;  // bucket counts, then starts, then ends
        for (int k = threadIdx.x; k < nbk; k += kBT) E[k] = 0;
        for (int k = threadIdx.x; k < kWin; k += kBT) bm[k] = 0u;  // each window re-zeroes what it used
        __syncthreads();
        stage_rows(q, nr, cap, long_ok, nlong, npieces, [&](int32_t x, bool ok) {
            run_add<false>(E, ok && x < span ? x >> sh : -1);
        });
        __syncthreads();
        {
            constexpr int kBPT = kMaxBuckets / kBT;
            int tid = (int)threadIdx.x;  // opaque, as above
            asm volatile("" : "+v"(tid));
            const int b0 = tid * kBPT;
            const int b1 = min(nbk, b0 + kBPT);
            int sum = 0;
            for (int k = b0; k < b1; ++k) sum += E[k];
            int64_t tt;
            int run = (int)block_excl_scan(sum, &tt);
            for (int k = b0; k < b1; ++k) {
                const int c = E[k];
                E[k] = run;
                run += c;
            }
        }
        __syncthreads();
        stage_rows(q, nr, cap, long_ok, nlong, npieces, [&](int32_t x, bool ok) {
            const bool in = ok && x < span;
            const int at = run_add<true>(E, in ? x >> sh : -1);
            if (in) buf[at] = x;
        });
        __syncthreads();
        // every entry, the F1c list's too, is in the slot now: level 2's
        // output is built in L.f1 and copied out once (a barrier after
        // global stores waits for them to complete)
        const bool copy_out = !q.out_lds;
        q.out_lds = true;
        // E[b] is now the end of bucket b.  Smallest bucket >= b0 holding an
        // entry (nbk if none):
        auto next_nonempty = [&](int b0) {
            const int st = b0 > 0 ? E[b0 - 1] : 0;
            int a = b0, z = nbk;
            while (a < z) {
                const int mid = (a + z) >> 1;
                if (E[mid] > st) z = mid; else a = mid + 1;
            }
            return a;
        };
        const int bpw = kBmBits >> sh;
        int cur = next_nonempty(0);
        int no = 0;
        while (cur < nbk) {
            const int wend = min(cur + bpw, nbk);
            const int e0 = cur > 0 ? E[cur - 1] : 0, e1 = E[wend - 1];
            const int32_t lo = cur << sh;
            const int nw = (int)((((int64_t)(wend - cur) << sh) + 31) >> 5);
            constexpr int kU = 4;  // loads in flight per thread
            for (int k0 = e0 + threadIdx.x; k0 < e1; k0 += kU * kBT) {
                int32_t xv[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) xv[u] = k0 + u * kBT < e1 ? buf[k0 + u * kBT] : q.root;
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int32_t x = xv[u];
                    if (x != q.root) atomicOr(&bm[(x - lo) >> 5], 1u << ((x - lo) & 31));
                }
            }
            __syncthreads();
            no = extract_bits_staged(q, bm, nw, lo, cap, no);
            if (no >= cap) break;
            cur = next_nonempty(wend);
        }
        __syncthreads();  // the output in L.f1 is complete
        if (copy_out) {
            int tid = (int)threadIdx.x;  // opaque, as above
            asm volatile("" : "+v"(tid));
            for (int k = tid; k < no; k += kBT) q.S[k] = L.f1[k];
            if (tid == 0) L.s_lds = 3;
            __syncthreads();
            q.out_lds = false;
        }
        return no;
    }
    int64_t lo64 = 0;
    while (lo64 < span) {
        const int32_t lo = (int32_t)lo64;
        const int32_t hi = (int32_t)min(lo64 + kBmBits, span);
        for (int k = threadIdx.x; k < kWin; k += kBT) bm[k] = 0u;
        __syncthreads();
        int32_t nxt = INT_MAX;  // smallest unconsumed id at or above hi
        if (q.level == 2) {
            // the F1c list's part of the window, found by a search in LDS
            const int c = L.cursor[0];
            int a = c, z = q.n1c;
            while (a < z) {
                const int mid = (a + z) >> 1;
                if (L.f1[mid] < hi) a = mid + 1; else z = mid;
            }
            for (int k = c + threadIdx.x; k < a; k += kBT) {
                const int32_t x = L.f1[k];
                atomicOr(&bm[(x - lo) >> 5], 1u << ((x - lo) & 31));
            }
            if (a < q.n1c) nxt = L.f1[a];
            __syncthreads();  // every thread has read cursor[0]
            if (threadIdx.x == 0) L.cursor[0] = a;
        }
        // a thread's rows kRB at a time: their offsets, then their next
        // entries, are loaded together before any row is walked
        constexpr int kRB = 2;
        constexpr int kWalk = 8;
        for (int r0 = threadIdx.x; r0 < nr; r0 += kBT * kRB) {
            g_cint* pp[kRB];
            int cc[kRB], ee[kRB];
            bool mine[kRB];
            int32_t x0[kRB];
#pragma unroll
            for (int k = 0; k < kRB; ++k) {
                const int r = r0 + k * kBT;
                pp[k] = GROW(q.g.fwd_adj);
                cc[k] = ee[k] = 0;
                mine[k] = false;
                if (r < nr && !is_f1_row(q, r)) {
                    int b;
                    graph_row(q, r, pp[k], b, ee[k]);
                    mine[k] = !(long_ok && ee[k] - b > kLongRow);
                    cc[k] = L.cursor[r];
                }
            }
#pragma unroll
            for (int k = 0; k < kRB; ++k)
                x0[k] = (mine[k] && cc[k] < ee[k]) ? pp[k][cc[k]] : INT_MAX;
#pragma unroll
            for (int k = 0; k < kRB; ++k) {
                if (!mine[k]) continue;
                int c = cc[k];
                int32_t x = x0[k];
                bool more = c < ee[k];
                while (more) {
                    if (x >= hi) { nxt = min(nxt, x); break; }
                    atomicOr(&bm[(x - lo) >> 5], 1u << ((x - lo) & 31));
                    ++c;
                    more = c < ee[k];
                    if (!more) break;
                    // the next kWalk entries loaded together: one memory
                    // round trip per kWalk entries of the walk
                    int32_t nx[kWalk];
#pragma unroll
                    for (int j = 0; j < kWalk; ++j) nx[j] = c + j < ee[k] ? pp[k][c + j] : INT_MAX;
                    x = nx[0];
#pragma unroll
                    for (int j = 1; j < kWalk; ++j) {
                        if (x >= hi) break;
                        atomicOr(&bm[(x - lo) >> 5], 1u << ((x - lo) & 31));
                        ++c;
                        x = nx[j];
                    }
                    more = c < ee[k];
                }
                L.cursor[r0 + k * kBT] = c;
            }
        }
        for (int s = wid; s < nlong; s += kWaves) {
            const int r = L.lrow[s];
            g_cint* p; int b, e;
            graph_row(q, r, p, b, e);
            const int c = L.cursor[r];
            const int z = wave_lower_bound(p, c, e, hi);
            for (int k = c + lane; k < z; k += 64) {
                const int32_t x = p[k];
                atomicOr(&bm[(x - lo) >> 5], 1u << ((x - lo) & 31));
            }
            if (lane == 0) {
                L.cursor[r] = z;
                if (z < e) nxt = min(nxt, p[z]);
            }
        }
        nxt = block_min(nxt);
        if (threadIdx.x == 0 && q.root >= lo && q.root < hi)
            bm[(q.root - lo) >> 5] &= ~(1u << ((q.root - lo) & 31));
        __syncthreads();
        extract_bits(q, bm, kWin, lo, cap);
        if (L.nout >= cap || nxt == INT_MAX) break;
        lo64 = nxt;  // skip the empty stretch
    }
    return L.nout;
}

// Builds S for `root` into out[] (global); returns |S|.  The two levels
// share one call site of collect, so it is inlined once.
__device__ __forceinline__ int two_hop(const CsrView& g, int32_t root, int cap, int32_t* out,
                                       const BfsStage& st) {
    Query q{g, root, 1, 0, false, 0, st, (g_int*)out, true};
    int n = 0;
#pragma nounroll
    for (int level = 1; level <= 2; ++level) {
        n = collect(q, cap);
        if (level == 1) {
            q.level = 2;
            q.out_lds = false;
            q.n1c = n;
            if (n == cap) {
                q.f1max = L.f1[cap - 1];
                g_cint* fo = GROW(g.fwd_off); g_cint* bo = GROW(g.bwd_off);
                int32_t mx = -1;
                if (fo[root + 1] > fo[root]) mx = max(mx, GROW(g.fwd_adj)[fo[root + 1] - 1]);
                if (bo[root + 1] > bo[root]) mx = max(mx, GROW(g.bwd_adj)[bo[root + 1] - 1]);
                q.check_f1 = mx > q.f1max;
            }
            if (threadIdx.x == 0) { L.n1c = n; L.t_level1 = wall_clock64(); }
            __syncthreads();
        }
    }
    return n;
}

__device__ __forceinline__ double pow2d(int k) { return u2d((uint64_t)(1023 + k) << 52); }

// Fixed-point image of a weight given the neighbourhood's scale.
__device__ __forceinline__ uint64_t nb_fixed(float w, double scale) {
    double q = (double)w * scale;
    return q < 0x1p50 ? (uint64_t)q : (uint64_t)1 << 50;
}

// k_bfs_root: S and its fixed-point CDF for each distinct root that the
// wavefront paths (k_bfs_small, k_bfs_mid) handed over in big_list.
// Stand-alone mode (only_roots non-null): block i builds S for
// only_roots[i] into S_buf + i*cap, its size into only_len[i] and, with
// ticks, the wall-clock ticks to the end of the 1-hop level, to the end, and
// the 2-hop candidate entries; no CDF.
__global__ __launch_bounds__(kBT, 4) void k_bfs_root(CsrView g, const float* __restrict__ w,
                                                  const int32_t* __restrict__ uroot,
                                                  const int32_t* __restrict__ big_list,
                                                  const int32_t* __restrict__ big_count, int cap,
                                                  int32_t* __restrict__ S_buf,
                                                  uint64_t* __restrict__ Q_buf,
                                                  int32_t* __restrict__ q_idx,
                                                  int32_t* __restrict__ ns_buf,
                                                  unsigned long long* __restrict__ bytes_out,
                                                  const int32_t* __restrict__ only_roots,
                                                  int32_t* __restrict__ only_len,
                                                  int64_t* __restrict__ ticks, BfsStage stage) {
    const bool only = only_roots != nullptr;
    if (!only && (int)blockIdx.x >= *big_count) return;
    const int64_t u = only ? (int64_t)blockIdx.x : big_list[blockIdx.x];
    const int32_t root = only ? only_roots[u] : uroot[u];
    int32_t* S = S_buf + u * cap;
    const bool valid = root >= 0 && root < g.n;
    if (threadIdx.x == 0) {
        L.slot = -1;
        L.t_start = wall_clock64();  // kept in LDS, not in a register across two_hop
    }
    const int ns = valid ? two_hop(g, root, cap, S, stage) : 0;
    if (threadIdx.x == 0 && L.slot >= 0) {
        // every access to the slot ended before two_hop's last barrier
        __threadfence();
        atomicExch(&stage.busy[L.slot], 0);
    }
    if (only) {
        if (threadIdx.x == 0) {
            only_len[u] = ns;
            if (ticks) {
                ticks[3 * u + 0] = valid ? L.t_level1 - L.t_start : 0;
                ticks[3 * u + 1] = wall_clock64() - L.t_start;
                ticks[3 * u + 2] = valid ? L.tot[2] : 0;
            }
        }
        return;
    }
    if (threadIdx.x == 0) {
        ns_buf[u] = ns;
        const int64_t b = 16 * (1 + (int64_t)L.n1c) + 4 * (L.tot[1] + L.tot[2] - L.n1c) +
                          4 * (int64_t)ns;
        if (bytes_out) atomicAdd(&bytes_out[(blockIdx.x % kByteSlots) * kByteSlotStride], (unsigned long long)b);
    }
    if (!w) return;  // S only (no weights given): no CDF
    __syncthreads();
    // fixed-point CDF: scale from the largest weight's exponent field.  Each
    // thread's (at most kMaxPer) weights are gathered once, loads issued
    // together, and kept in registers.
    constexpr int kMaxPer = kBfsCapMax / kBT;
    const int per = (ns + kBT - 1) / kBT;
    const int s0 = threadIdx.x * per, s1 = min(ns, s0 + per);
    int32_t xs[kMaxPer];
    float wv[kMaxPer];
    // S as the level left it in LDS, when it did (else the global copy)
    const int sl = L.s_lds;
    const int32_t* Sl = sl == 1 ? L.win : sl == 2 ? L.cursor : sl == 3 ? L.f1 : nullptr;
#pragma unroll
    for (int e = 0; e < kMaxPer; ++e) xs[e] = s0 + e < s1 ? (Sl ? Sl[s0 + e] : S[s0 + e]) : 0;
#pragma unroll
    for (int e = 0; e < kMaxPer; ++e) wv[e] = s0 + e < s1 ? w[xs[e]] : 0.0f;
    uint32_t mb = 0;
#pragma unroll
    for (int e = 0; e < kMaxPer; ++e) mb = max(mb, f2u(wv[e]));
    for (int off = 32; off > 0; off >>= 1) mb = max(mb, (uint32_t)__shfl_down((int)mb, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) L.red[threadIdx.x >> 6] = (int64_t)mb;
    __syncthreads();
    uint32_t gmb = 0;
    for (int k = 0; k < kWaves; ++k) gmb = max(gmb, (uint32_t)L.red[k]);
    const double scale = pow2d(49 - ((int)(gmb >> 23) - 127));
    int64_t acc = 0;
#pragma unroll
    for (int e = 0; e < kMaxPer; ++e) acc += s0 + e < s1 ? (int64_t)nb_fixed(wv[e], scale) : 0;
    int64_t tot;
    int64_t run = block_excl_scan(acc, &tot);
    uint64_t* Q = Q_buf + u * cap;
    int32_t* qidx = q_idx + u * kQIdx;
    const int qsh = q_shift((uint64_t)tot);
#pragma unroll
    for (int e = 0; e < kMaxPer; ++e) {
        if (s0 + e < s1) {
            const int64_t prev = run;
            run += (int64_t)nb_fixed(wv[e], scale);
            Q[s0 + e] = (uint64_t)run;
            q_index_put(qidx, s0 + e, ns, (uint64_t)prev, (uint64_t)run, qsh);
        }
    }
}


// ---- wavefront paths for small and medium neighbourhoods ----
//
// One 64-lane wavefront per distinct root whose 1-hop rows hold at most F1
// entries and whose 2-hop candidate list (F1c plus the rows of F1c) at most
// WIN entries.  The small tier (512, 2048; 10 KB of LDS, 16 roots per CU)
// takes every root and hands larger ones to the medium tier (1024, 4096;
// 20 KB, 8 per CU), which hands the rest to k_bfs_root.  Same result as the
// workgroup path: sort, drop duplicates (ballot/popcount compaction),
// filter, keep the cap smallest.
constexpr int kSmallF1 = 512;
constexpr int kSmallWin = kBfsSmallWinMax;
constexpr int kMidF1 = 1024;
constexpr int kMidWin = kBfsMidWinMax;

template <int F1, int WIN>
struct WaveLds {
    alignas(16) int32_t f1[F1];
    alignas(16) int32_t win[WIN];
};

// s[0,P), P a power of two <= WIN: padded to 64 if smaller (with INT_MAX),
// sorted.
template <int WIN>
__device__ __forceinline__ void wave_bitonic(int32_t* s, int P) {
    const int lane = threadIdx.x;
    if (P < 64) {
        if (lane >= P) s[lane] = INT_MAX;
        P = 64;
    }
    __syncthreads();
    switch (P) {
        case 64: wave_sort_reg<1>(s); break;
        case 128: wave_sort_reg<2>(s); break;
        case 256: wave_sort_reg<4>(s); break;
        case 512: wave_sort_reg<8>(s); break;
        case 1024: wave_sort_reg<16>(s); break;
        case 2048: wave_sort_reg<32>(s); break;
        default:
            if constexpr (WIN >= 4096) wave_sort_reg<64>(s);
            break;
    }
    __syncthreads();
}

// Ordered in-place compaction of the sorted s[0,n): keeps the first copy of
// each id for which keep(x) holds, at most `limit` of them; returns count.
template <typename Keep>
__device__ int wave_unique_compact(int32_t* s, int n, int limit, Keep keep) {
    const int lane = threadIdx.x;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int out = 0;
    int32_t carry = INT_MIN;  // last element of the previous chunk
    for (int base = 0; base < n && out < limit; base += 64) {
        const int i = base + lane;
        const int32_t x = i < n ? s[i] : INT_MAX;
        int32_t prev = __shfl_up(x, 1, 64);
        if (lane == 0) prev = carry;
        const bool k = i < n && x != prev && x != INT_MAX && keep(x);
        const uint64_t mask = __ballot(k);
        const int pos = out + __popcll(mask & below);
        carry = __shfl(x, 63, 64);
        __syncthreads();
        if (k && pos < limit) s[pos] = x;
        __syncthreads();
        out += __popcll(mask);
    }
    return out < limit ? out : limit;
}

// The body of both tiers.  FROM_LIST: block i takes root u = in_list[i]
// (i < *in_count) instead of u = i (< *n_unique_ptr).  Roots over the
// tier's limits are appended to out_list.
template <int F1, int WIN, bool FROM_LIST>
__device__ __forceinline__ void bfs_wave(CsrView g, const float* __restrict__ w,
                                         const int32_t* __restrict__ uroot,
                                         const int64_t* __restrict__ n_unique_ptr,
                                         const int32_t* __restrict__ in_list,
                                         const int32_t* __restrict__ in_count, int cap,
                                         int32_t* __restrict__ S_buf, uint64_t* __restrict__ Q_buf,
                                         int32_t* __restrict__ q_idx,
                                         int32_t* __restrict__ ns_buf,
                                         int32_t* __restrict__ out_list,
                                         int32_t* __restrict__ out_count,
                                         unsigned long long* __restrict__ bytes_out,
                                         int win_max) {
    __shared__ WaveLds<F1, WIN> Ls;
    int64_t u;
    if (FROM_LIST) {
        if ((int)blockIdx.x >= *in_count) return;
        u = in_list[blockIdx.x];
    } else {
        u = blockIdx.x;
        if (u >= *n_unique_ptr) return;
    }
    const int lane = threadIdx.x;
    const int32_t root = uroot[u];
    if (root < 0 || root >= g.n) {  // caller-given roots: out of range = empty S
        if (lane == 0) ns_buf[u] = 0;
        return;
    }
    const int32_t fb = g.fwd_off[root], fe = g.fwd_off[root + 1];
    const int32_t bb = g.bwd_off[root], be = g.bwd_off[root + 1];
    const int df = fe - fb, d1 = df + (be - bb);
    if (d1 > F1) {
        if (lane == 0) out_list[atomicAdd(out_count, 1)] = (int32_t)u;
        return;
    }
    // 1-hop: both rows, sorted, deduplicated, root dropped, capped
    int P1 = 1;
    while (P1 < d1) P1 <<= 1;
    for (int k = lane; k < P1; k += 64)
        Ls.f1[k] = k < df ? g.fwd_adj[fb + k] : (k < d1 ? g.bwd_adj[bb + k - df] : INT_MAX);
    __syncthreads();
    if (d1 > 1) wave_bitonic<F1>(Ls.f1, P1);
    const int n1 = wave_unique_compact(Ls.f1, d1, F1, [&](int32_t x) { return x != root; });
    const int n1c = n1 < cap ? n1 : cap;
    // 2-hop candidate count: the F1c rows' offsets, all loads of a lane
    // issued together
    constexpr int kChunks = F1 / 64;
    int64_t t = 0;
#pragma unroll
    for (int ch = 0; ch < kChunks; ++ch) {
        const int q = ch * 64 + lane;
        if (q < n1c) {
            const int32_t v = Ls.f1[q];
            t += (g.fwd_off[v + 1] - g.fwd_off[v]) + (g.bwd_off[v + 1] - g.bwd_off[v]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    const int64_t tot2 = t + n1c;
    if (tot2 > win_max) {
        if (lane == 0) out_list[atomicAdd(out_count, 1)] = (int32_t)u;
        return;
    }
    const bool check_f1 = n1 > n1c;
    const int32_t f1max = n1c > 0 ? Ls.f1[n1c - 1] : -1;
    // gather F1c and the rows of F1c, 64 row pairs at a time: lane q owns
    // row pair q (start `pos` in the chunk); the chunk's positions are filled
    // lane-strided, kGather at a time with their loads issued together, each
    // position finding its row pair by a binary search over the lanes'
    // starts, so consecutive lanes read consecutive row entries
    for (int q = lane; q < n1c; q += 64) Ls.win[q] = Ls.f1[q];
    int run = n1c;
    constexpr int kGather = 8;
    for (int q0 = 0; q0 < n1c; q0 += 64) {
        const int q = q0 + lane;
        int32_t a = 0, b = 0, c = 0, d = 0;
        if (q < n1c) {
            const int32_t v = Ls.f1[q];
            a = g.fwd_off[v]; b = g.fwd_off[v + 1]; c = g.bwd_off[v]; d = g.bwd_off[v + 1];
        }
        const int fl = b - a;
        const int len = fl + (d - c);
        int x = len;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        const int pos = x - len;                     // start within the chunk
        const int total = __shfl(x, 63, 64);
        for (int p0 = 0; p0 < total; p0 += 64 * kGather) {
            int32_t val[kGather];
#pragma unroll
            for (int e = 0; e < kGather; ++e) {
                const int pp = p0 + e * 64 + lane;
                // last lane whose start is <= pp (starts ascend across lanes)
                int r = 0;
#pragma unroll
                for (int st = 32; st > 0; st >>= 1) {
                    const int sp = __shfl(pos, r + st, 64);
                    if (sp <= pp) r += st;
                }
                const int k = pp - __shfl(pos, r, 64);
                const int32_t ar = __shfl(a, r, 64);
                const int flr = __shfl(fl, r, 64);
                const int32_t cr = __shfl(c, r, 64);
                val[e] = 0;
                if (pp < total) val[e] = k < flr ? g.fwd_adj[ar + k] : g.bwd_adj[cr + k - flr];
            }
#pragma unroll
            for (int e = 0; e < kGather; ++e) {
                const int pp = p0 + e * 64 + lane;
                if (pp < total) Ls.win[run + pp] = val[e];
            }
        }
        run += total;
    }
    int P2 = 1;
    while (P2 < (int)tot2) P2 <<= 1;
    for (int k = (int)tot2 + lane; k < P2; k += 64) Ls.win[k] = INT_MAX;
    __syncthreads();
    if (tot2 > 1) wave_bitonic<WIN>(Ls.win, P2);
    const int ns = wave_unique_compact(Ls.win, (int)tot2, cap, [&](int32_t x) {
        if (x == root) return false;
        if (check_f1 && x > f1max) {
            if (row_contains(GROW(g.fwd_adj), fb, fe, x)) return false;
            if (row_contains(GROW(g.bwd_adj), bb, be, x)) return false;
        }
        return true;
    });
    int32_t* S = S_buf + u * cap;
    if (!w) {  // S only (no weights given): no CDF
        for (int s2 = lane; s2 < ns; s2 += 64) S[s2] = Ls.win[s2];
        if (lane == 0) ns_buf[u] = ns;
        return;
    }
    // S, and its fixed-point CDF from the weights.  Pass 1 writes S and
    // gathers the weights (kWb per lane per round, loads issued together)
    // into the LDS slots of their ids; pass 2 reads four consecutive
    // weights per lane, sums them locally and scans the lane totals.
    uint64_t* Q = Q_buf + u * cap;
    constexpr int kWb = 16;
    uint32_t mb = 0;
    for (int s0 = 0; s0 < ns; s0 += 64 * kWb) {
        int32_t xs[kWb];
        float wb[kWb];
#pragma unroll
        for (int e = 0; e < kWb; ++e) {
            const int s2 = s0 + e * 64 + lane;
            xs[e] = s2 < ns ? Ls.win[s2] : 0;
        }
#pragma unroll
        for (int e = 0; e < kWb; ++e) wb[e] = s0 + e * 64 + lane < ns ? w[xs[e]] : 0.0f;
#pragma unroll
        for (int e = 0; e < kWb; ++e) {
            const int s2 = s0 + e * 64 + lane;
            if (s2 < ns) {
                S[s2] = xs[e];
                Ls.win[s2] = (int32_t)f2u(wb[e]);
                mb = max(mb, f2u(wb[e]));
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) mb = max(mb, (uint32_t)__shfl_xor((int)mb, off, 64));
    const double scale = pow2d(49 - ((int)(mb >> 23) - 127));
    // the CDF's total first (each lane reads back the weights it stored),
    // for the bucket index written with the CDF
    uint64_t zs = 0;
    for (int s2 = lane; s2 < ns; s2 += 64) zs += nb_fixed(__uint_as_float((uint32_t)Ls.win[s2]), scale);
    for (int off = 32; off > 0; off >>= 1) zs += __shfl_xor(zs, off, 64);
    const int qsh = q_shift(zs);
    int32_t* qidx = q_idx + u * kQIdx;
    __syncthreads();
    uint64_t carry = 0;
    for (int s0 = 0; s0 < ns; s0 += 256) {
        const int b0 = s0 + lane * 4;
        uint64_t q[4];
        uint64_t acc = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s2 = b0 + e;
            acc += s2 < ns ? nb_fixed(__uint_as_float((uint32_t)Ls.win[s2]), scale) : 0;
            q[e] = acc;
        }
        uint64_t x = acc;
        for (int off = 1; off < 64; off <<= 1) {
            uint64_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        const uint64_t ex = carry + x - acc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (b0 + e < ns) {
                Q[b0 + e] = ex + q[e];
                q_index_put(qidx, b0 + e, ns, e == 0 ? ex : ex + q[e - 1], ex + q[e], qsh);
            }
        }
        carry += __shfl(x, 63, 64);
    }
    if (lane == 0) {
        ns_buf[u] = ns;
        const int64_t by = 16 * (1 + (int64_t)n1c) + 4 * ((int64_t)d1 + tot2 - n1c) + 4 * (int64_t)ns;
        if (bytes_out) atomicAdd(&bytes_out[(blockIdx.x % kByteSlots) * kByteSlotStride], (unsigned long long)by);
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_bfs_small(
    CsrView g, const float* __restrict__ w, const int32_t* __restrict__ uroot,
    const int64_t* __restrict__ n_unique_ptr, int cap, int32_t* __restrict__ S_buf,
    uint64_t* __restrict__ Q_buf, int32_t* __restrict__ q_idx, int32_t* __restrict__ ns_buf,
    int32_t* __restrict__ out_list, int32_t* __restrict__ out_count,
    unsigned long long* __restrict__ bytes_out, int win_max) {
    bfs_wave<kSmallF1, kSmallWin, false>(g, w, uroot, n_unique_ptr, nullptr, nullptr, cap, S_buf,
                                         Q_buf, q_idx, ns_buf, out_list, out_count, bytes_out,
                                         win_max);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_bfs_mid(
    CsrView g, const float* __restrict__ w, const int32_t* __restrict__ uroot,
    const int32_t* __restrict__ in_list, const int32_t* __restrict__ in_count, int cap,
    int32_t* __restrict__ S_buf, uint64_t* __restrict__ Q_buf, int32_t* __restrict__ q_idx,
    int32_t* __restrict__ ns_buf, int32_t* __restrict__ out_list,
    int32_t* __restrict__ out_count, unsigned long long* __restrict__ bytes_out, int win_max) {
    bfs_wave<kMidF1, kMidWin, true>(g, w, uroot, nullptr, in_list, in_count, cap, S_buf, Q_buf,
                                    q_idx, ns_buf, out_list, out_count, bytes_out, win_max);
}

// ---- root deduplication ----
//
// A paper's root is its first PA pick (none without picks).  Each paper sets
// its root's bit in a bitmap over the paper ids (zero between years); the
// distinct roots are numbered in ascending id order by a prefix sum of the
// words' popcounts (so consecutive blocks of the BFS take nearby roots), and
// every paper reads its root's number from the bitmap.  A few passes over
// the year's papers and the N/32 words instead of a sort of the papers.

__device__ __forceinline__ int32_t paper_root(int64_t n_exist, int64_t i, int m,
                                              const int32_t* cit, const int32_t* cnt) {
    if (cnt[i] <= 0) return -1;
    const int32_t r = cit[i * m];
    return (r >= 0 && r < n_exist) ? r : -1;
}

__global__ void k_root_mark(int64_t n_exist, int64_t n_new, int m, const int32_t* __restrict__ cit,
                            const int32_t* __restrict__ cnt, uint32_t* __restrict__ bits) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new) return;
    const int32_t r = paper_root(n_exist, i, m, cit, cnt);
    if (r >= 0) atomicOr(&bits[r >> 5], 1u << (r & 31));
}

__global__ void k_root_popc(int64_t nw, const uint32_t* __restrict__ bits,
                            int32_t* __restrict__ wcount) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) wcount[w] = __popc(bits[w]);
}

__global__ void k_root_number(int64_t n_exist, int64_t n_new, int m, const int32_t* __restrict__ cit,
                              const int32_t* __restrict__ cnt, const uint32_t* __restrict__ bits,
                              const int64_t* __restrict__ wprefix, int64_t nw,
                              int64_t* __restrict__ n_unique, int32_t* __restrict__ root_slot,
                              int32_t* __restrict__ uroot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *n_unique = wprefix[nw];
    if (i >= n_new) return;
    const int32_t r = paper_root(n_exist, i, m, cit, cnt);
    if (r < 0) {
        root_slot[i] = -1;
        return;
    }
    const uint32_t below = bits[r >> 5] & ((1u << (r & 31)) - 1u);
    const int32_t u = (int32_t)(wprefix[r >> 5] + __popc(below));
    root_slot[i] = u;
    uroot[u] = r;  // every paper of the root writes the same value
}

__global__ void k_root_release(int64_t n_exist, int64_t n_new, int m,
                               const int32_t* __restrict__ cit, const int32_t* __restrict__ cnt,
                               uint32_t* __restrict__ bits) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new) return;
    const int32_t r = paper_root(n_exist, i, m, cit, cnt);
    if (r >= 0) bits[r >> 5] = 0u;
}

// ---- per-paper neighbourhood draws and refill ----

__device__ __forceinline__ bool row_has(const int32_t* row, int n, int32_t x) {
    for (int j = 0; j < n; ++j) if (row[j] == x) return true;
    return false;
}

// Neighbourhood draws and PA refill, kNbG lanes per paper.  After the PA
// phase the paper's row holds c0 picks (kept in LDS for the duplicate
// checks).  If S \ row has at most n_nb ids, all of them are taken in S
// order (lanes take consecutive S entries, a ballot keeps the order);
// otherwise draw attempts b = 0, 1, ... each pick S[upper_bound(Q, t)], and
// the lanes evaluate kNbG consecutive attempts at once: attempt b is taken
// iff its pick is neither in the row nor the pick of an earlier attempt of
// the same round -- the sequential redraw-on-duplicate rule.  The refill
// continues the paper's PA attempt sequence (one lane; it runs only when
// the neighbourhood could not supply n_nb ids).
// BATCH: the paper's simulation (batch table) supplies n_exist, n_nb, the
// CDF region and the Philox key; ids in rows and in S are global.
constexpr int kNbG = 16;  // measured against 4, 8 and 32 on the 2M-paper workload

template <bool BATCH>
__global__ __launch_bounds__(256) void k_nb_draw(Batch bt, int64_t n_exist, int64_t n_new, int m,
                                                 int n_nb, int cap, uint64_t seed, uint32_t yi,
                                                 const double* __restrict__ cdf,
                                                 const int32_t* __restrict__ cidx,
                                                 const int32_t* __restrict__ root_slot,
                                                 const int32_t* __restrict__ S_buf,
                                                 const uint64_t* __restrict__ Q_buf,
                                                 const int32_t* __restrict__ q_idx,
                                                 const int32_t* __restrict__ ns_buf,
                                                 int32_t* __restrict__ cit, int8_t* __restrict__ kind,
                                                 int32_t* __restrict__ cnt, int32_t* __restrict__ near,
                                                 int32_t* __restrict__ anext) {
    constexpr int G = kNbG;
    constexpr int kGroups = 256 / G;
    __shared__ int32_t s_row[kGroups][kMaxCites];
    __shared__ int32_t s_pick[256];
    const int grp = threadIdx.x / G;
    const int gl = threadIdx.x % G;
    const int lane = threadIdx.x & 63;
    const int shift = (lane / G) * G;
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << shift);
    const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t i = (int64_t)blockIdx.x * kGroups + grp;
    if (BATCH) n_new = bt.n_new;
    int64_t local = i, base = 0;
    if (BATCH && i < n_new) {
        const BatchSim& sm = bt.sims[batch_sim_of_new(bt, i)];
        n_exist = sm.n_exist;
        n_nb = sm.n_nb;
        base = sm.base;
        cdf += base;
        seed = sm.seed;
        local = i - sm.new_first;
    }
    const int c0 = (i < n_new && n_nb > 0) ? cnt[i] : 0;
    const bool live = c0 > 0;
    const uint32_t gid = (uint32_t)(n_exist + local);
    int32_t* row = cit + i * m;
    int8_t* kr = kind + i * m;
    int ns = 0;
    const int32_t* S = S_buf;
    const uint64_t* Q = Q_buf;
    const int32_t* qidx = q_idx;
    if (live) {
        const int64_t u = root_slot[i];
        ns = ns_buf[u];
        S = S_buf + u * cap;
        Q = Q_buf + u * cap;
        qidx = q_idx + u * kQIdx;
        for (int j = gl; j < c0; j += G) s_row[grp][j] = row[j];
    }
    // a paper's lanes (one wavefront) are all that read its LDS rows and
    // picks, so the rounds below synchronise wavefronts, not the block: a
    // wavefront stops with its own papers, not with the block's slowest
    wave_lds_sync();
    // PA picks inside S (only needed when taking all of S \ row is possible)
    int inS = 0;
    if (live && ns <= c0 + n_nb) {
        for (int j = gl; j < c0; j += G) {
            const int32_t x = s_row[grp][j];
            int lo = 0, hi = ns;
            while (lo < hi) { int mid = (lo + hi) >> 1; if (S[mid] < x) lo = mid + 1; else hi = mid; }
            inS += (lo < ns && S[lo] == x) ? 1 : 0;
        }
    }
    for (int off = 1; off < G; off <<= 1) inS += __shfl_xor(inS, off, 64);
    const bool take_all = live && ns - inS <= n_nb;
    int c = c0;
    // take all of S \ row, in S order
    bool act = take_all;
    int sb = 0;
    while (__ballot(act) != 0) {
        bool keep = false;
        int32_t x = 0;
        if (act) {
            const int sidx = sb + gl;
            if (sidx < ns) {
                x = S[sidx];
                keep = true;
                for (int j = 0; j < c0; ++j) keep &= s_row[grp][j] != x;
            }
        }
        const uint64_t km = __ballot(keep) & gmask;
        if (keep) {
            const int pos = c + __popcll(km & below);
            row[pos] = x;
            kr[pos] = 1;
            s_row[grp][pos] = x;
        }
        if (act) {
            c += __popcll(km);
            sb += G;
            if (sb >= ns) act = false;
        }
    }
    wave_lds_sync();  // the take-all rows are written
    // rejection draws over the neighbourhood's fixed-point CDF
    bool need = live && !take_all;
    const uint64_t Zq = (live && ns > 0) ? Q[ns - 1] : 0;
    const int qsh = q_shift(Zq);
    const int b_max = kAttemptsPerCite * n_nb;
    int got = 0, b0 = 0;
    while (__ballot(need) != 0) {
        const int b = b0 + gl;
        const bool valid = need && b < b_max;
        int32_t pick = -1;
        if (valid) {
            const u32x4 r = philox(gid, (uint32_t)b, yi, TAG_NB, seed);
            const uint64_t v = ((uint64_t)r.x << 21) ^ (uint64_t)(r.y >> 11);
            const uint64_t t = (__umul64hi(v, Zq) << 11) | ((v * Zq) >> 53);
            const int k = (int)(t >> qsh);  // t < Zq: k < kQBuckets
            int lo = qidx[k], hi = qidx[k + 1];
            while (lo < hi) { int mid = (lo + hi) >> 1; if (Q[mid] > t) hi = mid; else lo = mid + 1; }
            if (lo >= ns) lo = ns - 1;
            pick = S[lo];
        }
        s_pick[threadIdx.x] = pick;
        wave_lds_sync();
        bool dup = !valid;
        if (valid) {
            const int g0 = grp * G;
            for (int q = 0; q < gl && !dup; ++q) dup = s_pick[g0 + q] == pick;
            for (int q = 0; q < c && !dup; ++q) dup = s_row[grp][q] == pick;
        }
        const uint64_t acc = __ballot(!dup) & gmask;
        const int pos = got + __popcll(acc & below);
        const bool take = !dup && pos < n_nb;
        wave_lds_sync();
        if (take) {
            row[c + pos - got] = pick;
            kr[c + pos - got] = 1;
            s_row[grp][c + pos - got] = pick;
        }
        if (need) {
            const int add = min(__popcll(acc), n_nb - got);
            got += add;
            c += add;
            b0 += G;
            if (got >= n_nb || b0 >= b_max) need = false;
        }
    }
    wave_lds_sync();  // the drawn rows are written
    // refill with PA draws, continuing the attempt sequence
    if (live && gl == 0) {
        const double Zg = cdf[n_exist - 1];
        const double eps = kNearRel * Zg;
        const double kscale = cdf_key_scale(Zg);
        const int a_max = kAttemptsPerCite * m;
        int a = anext[i];
        int nr = 0;
        while (c < m && a < a_max) {
            u32x4 r = philox(gid, (uint32_t)a, yi, TAG_PA, seed);
            const double t = u53(r.x, r.y) * Zg;
            const int64_t k = cdf_search(cdf, cidx, n_exist, t, kscale);
            if (cdf_near(cdf, n_exist, k, t, eps)) ++nr;
            ++a;
            const int32_t x = (int32_t)(k + base);
            if (!row_has(s_row[grp], c, x)) {
                row[c] = x;
                kr[c] = 2;
                s_row[grp][c] = x;
                ++c;
            }
        }
        cnt[i] = c;
        near[i] += nr;
        anext[i] = a;
    }
}

// The 2-hop kernels' algorithmic-byte counts, spread over kByteSlots
// counters a cache line apart (one same-address atomic per root had queued
// tens of thousands of atomics on one L2 channel), summed into *out and
// cleared for the next launch.
__global__ __launch_bounds__(64) void k_fold_bytes(unsigned long long* __restrict__ slots,
                                                   unsigned long long* __restrict__ out) {
    static_assert(kByteSlots == 64, "one lane per counter");
    unsigned long long v = slots[threadIdx.x * kByteSlotStride];
    slots[threadIdx.x * kByteSlotStride] = 0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (threadIdx.x == 0) atomicAdd(out, v);
}

}  // namespace

void launch_root_dedup(hipStream_t s, int64_t n_exist, int64_t n_new, int m, const int32_t* cit,
                       const int32_t* cnt, uint32_t* root_bits, int32_t* wcount, int64_t* wprefix,
                       int64_t* scratch, int64_t* n_unique, int32_t* root_slot, int32_t* uroot) {
    const int64_t nw = (n_exist + 31) / 32;
    const int64_t nb = (std::max<int64_t>(n_new, 1) + 255) / 256;
    if (n_new > 0)
        hipLaunchKernelGGL(k_root_mark, dim3((uint32_t)nb), dim3(256), 0, s, n_exist, n_new, m, cit,
                           cnt, root_bits);
    hipLaunchKernelGGL(k_root_popc, dim3((uint32_t)((nw + 255) / 256)), dim3(256), 0, s, nw,
                       root_bits, wcount);
    launch_excl_scan_i32(s, wcount, nw, wprefix, scratch);
    hipLaunchKernelGGL(k_root_number, dim3((uint32_t)nb), dim3(256), 0, s, n_exist, n_new, m, cit,
                       cnt, root_bits, wprefix, nw, n_unique, root_slot, uroot);
    if (n_new > 0)
        hipLaunchKernelGGL(k_root_release, dim3((uint32_t)nb), dim3(256), 0, s, n_exist, n_new, m,
                           cit, cnt, root_bits);
}

void launch_bfs_root(hipStream_t s, const CsrView& g, const float* w, const int32_t* uroot,
                     const int64_t* n_unique, int64_t max_unique, int cap, int32_t* S_buf,
                     uint64_t* Q_buf, int32_t* q_idx, int32_t* ns_buf, int32_t* big_list,
                     int32_t* big_count, unsigned long long* bytes_out, const BfsStage& stage,
                     unsigned long long* byte_slots) {
    // the kernels count into the spread slots, folded into bytes_out after
    unsigned long long* const out = bytes_out;
    bytes_out = out ? byte_slots : nullptr;
    if (max_unique <= 0) return;
    int32_t* mid_list = big_list + max_unique;
    const int small_max = stage.small_max > 0 ? std::min(stage.small_max, kSmallWin) : kSmallWin;
    const int mid_max = stage.mid_max > 0 ? std::min(stage.mid_max, kMidWin) : kMidWin;
    hipLaunchKernelGGL(k_bfs_small, dim3((uint32_t)max_unique), dim3(64), 0, s, g, w, uroot,
                       n_unique, cap, S_buf, Q_buf, q_idx, ns_buf, big_list, big_count, bytes_out,
                       small_max);
    hipLaunchKernelGGL(k_bfs_mid, dim3((uint32_t)max_unique), dim3(64), 0, s, g, w, uroot,
                       big_list, big_count, cap, S_buf, Q_buf, q_idx, ns_buf, mid_list,
                       big_count + 1, bytes_out, mid_max);
    hipLaunchKernelGGL(k_bfs_root, dim3((uint32_t)max_unique), dim3(kBT), 0, s, g, w, uroot,
                       mid_list, big_count + 1, cap, S_buf, Q_buf, q_idx, ns_buf, bytes_out,
                       (const int32_t*)nullptr, (int32_t*)nullptr, (int64_t*)nullptr, stage);
    if (out && byte_slots) hipLaunchKernelGGL(k_fold_bytes, dim3(1), dim3(64), 0, s, byte_slots, out);
}

void launch_nb_draw(hipStream_t s, int64_t n_exist, int64_t n_new, int m, int n_nb, int cap,
                    uint64_t seed, uint32_t year_index, const double* cdf, const int32_t* cdf_idx,
                    const int32_t* root_slot, const int32_t* S_buf, const uint64_t* Q_buf,
                    const int32_t* q_idx, const int32_t* ns_buf, int32_t* cit, int8_t* kind,
                    int32_t* cnt, int32_t* near, int32_t* anext) {
    if (n_new <= 0 || n_nb <= 0) return;
    const int64_t nb = (n_new + 256 / kNbG - 1) / (256 / kNbG);
    hipLaunchKernelGGL(k_nb_draw<false>, dim3((uint32_t)nb), dim3(256), 0, s, Batch{}, n_exist,
                       n_new, m, n_nb, cap, seed, year_index, cdf, cdf_idx, root_slot, S_buf,
                       Q_buf, q_idx, ns_buf, cit, kind, cnt, near, anext);
}

void launch_nb_draw_batch(hipStream_t s, const Batch& b, int m, int cap, uint32_t year_index,
                          const double* cdf, const int32_t* root_slot, const int32_t* S_buf,
                          const uint64_t* Q_buf, const int32_t* q_idx, const int32_t* ns_buf,
                          int32_t* cit, int8_t* kind, int32_t* cnt, int32_t* near,
                          int32_t* anext) {
    if (b.n_new <= 0) return;
    const int64_t nb = (b.n_new + 256 / kNbG - 1) / (256 / kNbG);
    hipLaunchKernelGGL(k_nb_draw<true>, dim3((uint32_t)nb), dim3(256), 0, s, b, (int64_t)0,
                       b.n_new, m, 0, cap, (uint64_t)0, year_index, cdf, (const int32_t*)nullptr,
                       root_slot, S_buf, Q_buf, q_idx, ns_buf, cit, kind, cnt, near, anext);
}

BfsStage bfs_stage_view(const BfsStage& pool, int mode, int small_max, int mid_max) {
    BfsStage v = pool;
    v.max_total = mode == 1 ? 0 : mode == 2 ? INT64_MAX : kBfsStageDefaultMax;
    v.radix = mode == 0 ? 1 : 0;
    v.small_max = small_max;
    v.mid_max = mid_max;
    return v;
}

int bfs_stage_slots(int device) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs_root, kBT, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return 2 * std::max(per_cu, 1) * std::max(cus, 1);
}

void launch_bfs_only(hipStream_t s, const CsrView& g, int64_t count, const int32_t* roots, int cap,
                     int32_t* out, int32_t* out_len, int64_t* ticks, const BfsStage& stage) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_bfs_root, dim3((uint32_t)count), dim3(kBT), 0, s, g, (const float*)nullptr,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, (const int32_t*)nullptr,
                       cap, out, (uint64_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                       (unsigned long long*)nullptr, roots, out_len, ticks, stage);
}

}  // namespace pg
