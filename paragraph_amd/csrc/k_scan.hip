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
// sequential across the four wavefronts); tiles take tickets in launch order
// and look back to the nearest tile that has published its inclusive
// prefix, then add the intervening tile aggregates *forward*.  Forward
// addition from any published inclusive prefix reproduces the same
// left-to-right chain of tile sums, so the result does not depend on timing.
#include "pg_internal.h"

namespace pg {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = kTile / kScanThreads;  // 16

__device__ __forceinline__ uint32_t ld_acquire(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back over the 64 nearest predecessor tiles (wave 0 only): waits until
// one of them has published its inclusive prefix and every tile after it
// its aggregate, then leaves the start value and those aggregates in LDS.
struct LookBack {
    double vals[64];
    double start;
    int from;
};

__device__ void lookback_collect(const ScanState st, uint32_t vb, LookBack* lb) {
    const int lane = threadIdx.x & 63;
    const uint32_t want_agg = st.epoch * 4u + 1u;
    const uint32_t want_inc = st.epoch * 4u + 2u;
    for (;;) {
        int64_t k = (int64_t)vb - 1 - lane;
        uint32_t stt = k >= 0 ? ld_acquire(st.status + k) : want_inc;
        bool inc = (stt == want_inc);
        bool ready = inc || stt == want_agg;
        uint64_t inc_mask = __ballot(inc);
        uint64_t rdy_mask = __ballot(ready);
        int first = inc_mask ? __builtin_ctzll(inc_mask) : 64;
        uint64_t need = (first >= 64) ? ~0ull : ((1ull << first) - 1ull);
        if (first < 64 && (rdy_mask & need) == need) {
            if (lane < first) lb->vals[lane] = st.agg[vb - 1 - lane];
            if (lane == first) {
                lb->start = (k >= 0) ? st.incl[k] : 0.0;
                lb->from = first;
            }
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__global__ __launch_bounds__(kScanThreads) void k_weights_cdf(
    int64_t n, const int32_t* __restrict__ in_deg, const int32_t* __restrict__ birth,
    const float* __restrict__ logfit, int32_t year, float alpha, float inv_tau,
    float* __restrict__ w_out, double* __restrict__ cdf_out, ScanState st) {
    __shared__ uint32_t s_vb;
    __shared__ double s_wave[4];
    __shared__ double s_excl;
    if (threadIdx.x == 0) s_vb = atomicAdd(st.ticket, 1u);
    __syncthreads();
    const uint32_t vb = s_vb;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int64_t base = (int64_t)vb * kTile + (int64_t)threadIdx.x * kScanItems;

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
    double s[kScanItems];
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) { acc += (double)wv[j]; s[j] = acc; }
    // Kogge-Stone inclusive scan of thread totals across the wavefront.
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
    double v[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) v[j] = texcl + s[j];

    // Publish this tile's aggregate (value at its last element).
    if (threadIdx.x == kScanThreads - 1) {
        if (vb == 0) {
            st.incl[0] = v[kScanItems - 1];
            __threadfence();
            st_release(st.status + 0, st.epoch * 4u + 2u);
        } else {
            st.agg[vb] = v[kScanItems - 1];
            __threadfence();
            st_release(st.status + vb, st.epoch * 4u + 1u);
        }
    }
    __shared__ LookBack s_lb;
    if (vb != 0 && wid == 0) lookback_collect(st, vb, &s_lb);
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0;
        if (vb != 0) {
            e = s_lb.start;
            for (int q = s_lb.from - 1; q >= 0; --q) e += s_lb.vals[q];  // oldest first
        }
        s_excl = e;
    }
    __syncthreads();
    const double e = s_excl;
    if (vb != 0 && threadIdx.x == kScanThreads - 1) {
        st.incl[vb] = e + v[kScanItems - 1];
        __threadfence();
        st_release(st.status + vb, st.epoch * 4u + 2u);
    }
    if (base + kScanItems <= n) {
        double2* co = reinterpret_cast<double2*>(cdf_out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 2; ++q)
            co[q] = make_double2(e + v[2 * q], e + v[2 * q + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j)
            if (base + j < n) cdf_out[base + j] = e + v[j];
    }
    if (threadIdx.x == 0 && vb == gridDim.x - 1) atomicExch(st.ticket, 0u);
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

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_sum_i32(const int32_t* __restrict__ p, int64_t n, unsigned long long* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long acc = 0;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) acc += (unsigned long long)(int64_t)p[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

}  // namespace

void launch_weights_cdf(hipStream_t s, int64_t n, const int32_t* in_deg, const int32_t* birth,
                        const float* logfit, int32_t year, float alpha, float inv_tau, float* w,
                        double* cdf, const ScanState& st) {
    if (n <= 0) return;
    int64_t nb = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_weights_cdf, dim3((uint32_t)nb), dim3(kScanThreads), 0, s, n, in_deg,
                       birth, logfit, year, alpha, inv_tau, w, cdf, st);
}

int64_t scan_scratch_elems(int64_t n) { return (n + kTile - 1) / kTile + 2; }

void launch_excl_scan_i32(hipStream_t s, const int32_t* in, int64_t n, int64_t* out,
                          int64_t* scratch) {
    int64_t nt = (n + kTile - 1) / kTile;
    if (nt == 0) nt = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nt), dim3(256), 0, s, in, n, scratch);
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
    int64_t nb = (n + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_sum_i32, dim3((uint32_t)nb), dim3(256), 0, s, p, n, out);
}

}  // namespace pg
