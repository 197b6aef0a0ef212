// k_scan.hip -- fused PA weight pass + single-pass FP64 CDF, and integer
// scans.
//
// Replaces the reference's host score preparation (FillInDegreeArr,
// FillFitnessArr, CalculateScores: src/abm.cu:112-176, src/abm.cu:324-338,
// called per year at src/kernel.cu:1698-1706) and the per-year H2D uploads
// of the score arrays (src/kernel.cu:1177-1191).  The state stays in HBM;
// one launch reads in_deg/birth/logfit (12 B per paper) and writes the FP32
// weight and the FP64 CDF.
//
// The CDF is deterministic: each 4096-element tile is summed in a fixed
// order (sequential within a thread, Kogge-Stone across a wavefront,
// sequential across the four wavefronts), the tile totals are chained left
// to right, and every element is its tile's exclusive prefix plus its local
// value.  Three launches (tile pass, tile-total chain, down-sweep) keep
// inter-block communication at kernel boundaries: a single-pass look-back
// needs a device-scope release per tile, which on this part writes back L2
// and serialises the tiles.
#include "pg_internal.h"

namespace pg {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = kTile / kScanThreads;  // 16

// Local (in-tile) inclusive values of one thread's 16 weights, in the fixed
// summation order; *tile_last gets the tile total for thread 255.
struct TileScan {
    double v[kScanItems];
};

__device__ __forceinline__ TileScan tile_local_scan(const float* wv, double* s_wave) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    double s[kScanItems];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) { acc += (double)wv[j]; s[j] = acc; }
    double x = acc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double y = __shfl_up(x, off, 64);
        if (lane >= off) x = y + x;
    }
    double ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0.0;
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    double wpre = 0.0;
    for (int k = 0; k < wid; ++k) wpre += s_wave[k];
    const double texcl = (wid == 0) ? ex : wpre + ex;
    TileScan t;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) t.v[j] = texcl + s[j];
    return t;
}

// Tile of this block: plain (tile blockIdx.x of [0, n)) or batched (the
// simulation holding batch tile blockIdx.x; n becomes its region end and
// alpha/inv_tau its own).
struct TileAt {
    int64_t start, end;
    float alpha, inv_tau;
};

template <bool BATCH>
__device__ __forceinline__ TileAt tile_at(const Batch& b, int64_t n, float alpha, float inv_tau) {
    TileAt t;
    if (BATCH) {
        const BatchSim& sm = b.sims[batch_sim_of_tile(b, blockIdx.x)];
        t.start = sm.base + ((int64_t)blockIdx.x - sm.tile_first) * kTile;
        t.end = sm.base + sm.n_exist;
        t.alpha = sm.alpha;
        t.inv_tau = sm.inv_tau;
    } else {
        t.start = (int64_t)blockIdx.x * kTile;
        t.end = n;
        t.alpha = alpha;
        t.inv_tau = inv_tau;
    }
    return t;
}

// Pass 1: weights (stored) and each tile's total.
template <bool BATCH>
__global__ __launch_bounds__(kScanThreads) void k_weights_tiles(
    Batch bt, int64_t n_plain, const int32_t* __restrict__ in_deg,
    const int32_t* __restrict__ birth, const float* __restrict__ logfit, int32_t year,
    float alpha_plain, float inv_tau_plain, float* __restrict__ w_out,
    double* __restrict__ tile_total) {
    __shared__ double s_wave[4];
    const TileAt ta = tile_at<BATCH>(bt, n_plain, alpha_plain, inv_tau_plain);
    const int64_t n = ta.end;
    const float alpha = ta.alpha, inv_tau = ta.inv_tau;
    const int64_t base = ta.start + (int64_t)threadIdx.x * kScanItems;
    float wv[kScanItems];
    if (base + kScanItems <= n) {
        const int4* kd = reinterpret_cast<const int4*>(in_deg + base);
        const int4* bd = reinterpret_cast<const int4*>(birth + base);
        const float4* fd = reinterpret_cast<const float4*>(logfit + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            int4 k4 = kd[q];
            int4 b4 = bd[q];
            float4 f4 = fd[q];
            wv[4 * q + 0] = pa_weight(k4.x, year - b4.x, f4.x, alpha, inv_tau);
            wv[4 * q + 1] = pa_weight(k4.y, year - b4.y, f4.y, alpha, inv_tau);
            wv[4 * q + 2] = pa_weight(k4.z, year - b4.z, f4.z, alpha, inv_tau);
            wv[4 * q + 3] = pa_weight(k4.w, year - b4.w, f4.w, alpha, inv_tau);
        }
        float4* wo = reinterpret_cast<float4*>(w_out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q)
            wo[q] = make_float4(wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            int64_t i = base + j;
            wv[j] = 0.0f;
            if (i < n) {
                wv[j] = pa_weight(in_deg[i], year - birth[i], logfit[i], alpha, inv_tau);
                w_out[i] = wv[j];
            }
        }
    }
    TileScan t = tile_local_scan(wv, s_wave);
    if (threadIdx.x == kScanThreads - 1) tile_total[blockIdx.x] = t.v[kScanItems - 1];
}

// Pass 2: exclusive prefix of the tile totals, chained left to right by one
// lane from LDS (the loads and stores of each 4096-tile chunk are spread
// over the block).  Batched: block s chains simulation s's tiles.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_tile_chain(Batch bt, double* __restrict__ tile_total,
                                                    int64_t ntiles) {
    __shared__ double s_t[kTile];
    __shared__ double s_carry;
    if (BATCH) {
        const BatchSim& sm = bt.sims[blockIdx.x];
        tile_total += sm.tile_first;
        ntiles = (sm.n_exist + kTile - 1) / kTile;
    }
    if (threadIdx.x == 0) s_carry = 0.0;
    for (int64_t c0 = 0; c0 < ntiles; c0 += kTile) {
        const int64_t cn = (ntiles - c0 < kTile) ? ntiles - c0 : kTile;
        for (int64_t k = threadIdx.x; k < cn; k += 256) s_t[k] = tile_total[c0 + k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double e = s_carry;
            for (int64_t k = 0; k < cn; ++k) {
                const double t = s_t[k];
                s_t[k] = e;
                e = e + t;
            }
            s_carry = e;
        }
        __syncthreads();
        for (int64_t k = threadIdx.x; k < cn; k += 256) tile_total[c0 + k] = s_t[k];
        __syncthreads();
    }
}

// Pass 3: every element's CDF value (tile prefix + local value).
template <bool BATCH>
__global__ __launch_bounds__(kScanThreads) void k_cdf_down(Batch bt, int64_t n_plain,
                                                          const float* __restrict__ w,
                                                          const double* __restrict__ tile_excl,
                                                          double* __restrict__ cdf_out) {
    __shared__ double s_wave[4];
    const TileAt ta = tile_at<BATCH>(bt, n_plain, 0.0f, 0.0f);
    const int64_t n = ta.end;
    const int64_t base = ta.start + (int64_t)threadIdx.x * kScanItems;
    float wv[kScanItems];
    if (base + kScanItems <= n) {
        const float4* wi = reinterpret_cast<const float4*>(w + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            float4 f = wi[q];
            wv[4 * q] = f.x; wv[4 * q + 1] = f.y; wv[4 * q + 2] = f.z; wv[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) wv[j] = base + j < n ? w[base + j] : 0.0f;
    }
    TileScan t = tile_local_scan(wv, s_wave);
    const double e = tile_excl[blockIdx.x];
    if (base + kScanItems <= n) {
        double2* co = reinterpret_cast<double2*>(cdf_out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 2; ++q)
            co[q] = make_double2(e + t.v[2 * q], e + t.v[2 * q + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j)
            if (base + j < n) cdf_out[base + j] = e + t.v[j];
    }
}

// ---- integer exclusive scan (reduce, scan of tile sums, down-sweep) ----

__device__ __forceinline__ int64_t block_excl_scan_i64(int64_t v, int64_t* total) {
    __shared__ int64_t s_w[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int k = 0; k < 4; ++k) { if (k < wid) pre += s_w[k]; tot += s_w[k]; }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                     int64_t* __restrict__ tile_sum) {
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int64_t acc = 0;
    for (int j = threadIdx.x; j < kTile; j += 256) {
        int64_t i = base + j;
        if (i < n) acc += in[i];
    }
    int64_t tot;
    block_excl_scan_i64(acc, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_tiles(int64_t* __restrict__ tile_sum, int64_t ntiles) {
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < ntiles; b0 += 256) {
        int64_t i = b0 + threadIdx.x;
        int64_t v = i < ntiles ? tile_sum[i] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan_i64(v, &tot);
        int64_t carry = s_carry;
        if (i < ntiles) tile_sum[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[ntiles] = s_carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const int32_t* __restrict__ in, int64_t n,
                                                   const int64_t* __restrict__ tile_off,
                                                   int64_t* __restrict__ out) {
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * 16;
    int32_t vals[16];
    int64_t acc = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int64_t i = base + j;
        vals[j] = i < n ? in[i] : 0;
        acc += vals[j];
    }
    int64_t tot;
    int64_t ex = block_excl_scan_i64(acc, &tot);
    int64_t run = tile_off[blockIdx.x] + ex;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int64_t i = base + j;
        if (i < n) out[i] = run;
        run += vals[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_off[gridDim.x];
}

// k_scan_tiles and k_scan_down in one launch for at most 256 tiles: each
// block sums the tile sums before its own (one per thread) first.
__global__ __launch_bounds__(256) void k_scan_down_fused(const int32_t* __restrict__ in, int64_t n,
                                                         const int64_t* __restrict__ tile_sum,
                                                         int64_t* __restrict__ out) {
    int64_t off;
    block_excl_scan_i64((int64_t)threadIdx.x < (int64_t)blockIdx.x ? tile_sum[threadIdx.x] : 0,
                        &off);
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * 16;
    int32_t vals[16];
    int64_t acc = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int64_t i = base + j;
        vals[j] = i < n ? in[i] : 0;
        acc += vals[j];
    }
    int64_t tot;
    int64_t ex = block_excl_scan_i64(acc, &tot);
    int64_t run = off + ex;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int64_t i = base + j;
        if (i < n) out[i] = run;
        run += vals[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Sum into *out: one atomic per block (a wave-level atomic per 64 lanes
// queued thousands of atomics on one address).
__global__ __launch_bounds__(256) void k_sum_i32(const int32_t* __restrict__ p, int64_t n,
                                                 unsigned long long* out) {
    __shared__ unsigned long long part[4];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long acc = 0;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) acc += (unsigned long long)(int64_t)p[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

// Bucket index of a CDF (see cdf_search): idx[k] = the first paper i with
// bucket(cdf[i]) >= k, n if none (k = 0 .. kCdfIdx + 1).  One thread per
// bucket, a binary search over the CDF for its boundary: a heavy paper spans
// many buckets, and a thread per paper writing its span made that paper's
// thread the launch's tail.
__global__ __launch_bounds__(256) void k_cdf_index(const double* __restrict__ cdf, int64_t n,
                                                   int32_t* __restrict__ idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > kCdfIdx + 1) return;
    const double kscale = cdf_key_scale(cdf[n - 1]);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf_key(cdf[mid], kscale) >= k) hi = mid; else lo = mid + 1;
    }
    idx[k] = (int32_t)lo;
}

__global__ void k_acc_i64(const int64_t* __restrict__ src, unsigned long long* dst) {
    if (threadIdx.x == 0) *dst += (unsigned long long)*src;
}

}  // namespace

void launch_acc_i64(hipStream_t s, const int64_t* src, unsigned long long* dst) {
    hipLaunchKernelGGL(k_acc_i64, dim3(1), dim3(64), 0, s, src, dst);
}

void launch_weights_cdf(hipStream_t s, int64_t n, const int32_t* in_deg, const int32_t* birth,
                        const float* logfit, int32_t year, float alpha, float inv_tau, float* w,
                        double* cdf, double* tile_scratch, int32_t* cdf_idx) {
    if (n <= 0) return;
    int64_t nb = (n + kTile - 1) / kTile;
    const Batch none{};
    hipLaunchKernelGGL(k_weights_tiles<false>, dim3((uint32_t)nb), dim3(kScanThreads), 0, s, none,
                       n, in_deg, birth, logfit, year, alpha, inv_tau, w, tile_scratch);
    hipLaunchKernelGGL(k_tile_chain<false>, dim3(1), dim3(256), 0, s, none, tile_scratch, nb);
    hipLaunchKernelGGL(k_cdf_down<false>, dim3((uint32_t)nb), dim3(kScanThreads), 0, s, none, n, w,
                       tile_scratch, cdf);
    if (cdf_idx)
        hipLaunchKernelGGL(k_cdf_index, dim3((uint32_t)((kCdfIdx + 2 + 255) / 256)), dim3(256), 0, s,
                           cdf, n, cdf_idx);
}

void launch_weights_cdf_batch(hipStream_t s, const Batch& b, const int32_t* in_deg,
                              const int32_t* birth, const float* logfit, float* w, double* cdf,
                              double* tile_scratch) {
    if (b.n_tiles <= 0) return;
    hipLaunchKernelGGL(k_weights_tiles<true>, dim3((uint32_t)b.n_tiles), dim3(kScanThreads), 0, s,
                       b, (int64_t)0, in_deg, birth, logfit, b.year, 0.0f, 0.0f, w, tile_scratch);
    hipLaunchKernelGGL(k_tile_chain<true>, dim3((uint32_t)b.n_sims), dim3(256), 0, s, b,
                       tile_scratch, (int64_t)0);
    hipLaunchKernelGGL(k_cdf_down<true>, dim3((uint32_t)b.n_tiles), dim3(kScanThreads), 0, s, b,
                       (int64_t)0, w, tile_scratch, cdf);
}

int64_t scan_scratch_elems(int64_t n) { return (n + kTile - 1) / kTile + 2; }

void launch_excl_scan_i32(hipStream_t s, const int32_t* in, int64_t n, int64_t* out,
                          int64_t* scratch) {
    int64_t nt = (n + kTile - 1) / kTile;
    if (nt == 0) nt = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nt), dim3(256), 0, s, in, n, scratch);
    if (nt <= 256) {
        hipLaunchKernelGGL(k_scan_down_fused, dim3((uint32_t)nt), dim3(256), 0, s, in, n, scratch,
                           out);
        return;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, scratch, nt);
    hipLaunchKernelGGL(k_scan_down, dim3((uint32_t)nt), dim3(256), 0, s, in, n, scratch, out);
}

void launch_fill_i32(hipStream_t s, int32_t* p, int32_t v, int64_t n) {
    if (n <= 0) return;
    int64_t nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3((uint32_t)nb), dim3(256), 0, s, p, v, n);
}

void launch_sum_i32(hipStream_t s, const int32_t* p, int64_t n, unsigned long long* out) {
    if (n <= 0) return;
    int64_t nb = (n + 1023) / 1024;
    if (nb > 256) nb = 256;
    hipLaunchKernelGGL(k_sum_i32, dim3((uint32_t)nb), dim3(256), 0, s, p, n, out);
}

}  // namespace pg
