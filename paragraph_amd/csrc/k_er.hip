// k_er.hip -- Erdos-Renyi G(n,p) edge formation for a year of new papers.
//
// The reference's ER model lives on the gpu-er branch, which is not
// available; its command line (--model er-gnp, --er-probability:
// results/slurm-8640738.out:553-556) and run logs (results/slurm-8634746.out:73)
// show each new paper linking to each existing paper independently with
// probability p.  Targets are produced in ascending order by geometric
// skips (Batagelj & Brandes, Phys. Rev. E 71, 036113, 2005):
//   gap = floor(log(u) / log1p(-p)),  u in (0,1] from Philox
//   counter (paper, skip index, year index, TAG_ER),
// with the deterministic FP64 log, so the edge set is bit-exact with the
// oracle.  Two passes (count, then write at scanned offsets) keep the
// output contiguous and ordered without atomics on positions.
#include "pg_internal.h"

namespace pg {
namespace {

template <bool kWrite>
__device__ __forceinline__ int64_t er_paper(int64_t n_exist, int64_t i, double p, double c,
                                            uint64_t seed, uint32_t yi, int32_t* out) {
    if (p <= 0.0) return 0;
    const uint32_t gid = (uint32_t)(n_exist + i);
    if (p >= 1.0) {
        if (kWrite) {
            for (int64_t t = 0; t < n_exist; ++t) out[t] = (int32_t)t;
        }
        return n_exist;
    }
    double pos = -1.0;
    int64_t e = 0;
    for (uint32_t j = 0;; ++j) {
        u32x4 r = philox(gid, j, yi, TAG_ER, seed);
        double u = u53_open0(r.x, r.y);
        double gap = __builtin_floor(det_log(u) / c);
        double nxt = pos + 1.0 + gap;
        if (nxt >= (double)n_exist) break;
        pos = nxt;
        if (kWrite) out[e] = (int32_t)pos;
        ++e;
    }
    return e;
}

__global__ __launch_bounds__(256) void k_er_count(int64_t n_exist, int64_t n_new, double p, double c,
                                                  uint64_t seed, uint32_t yi,
                                                  int32_t* __restrict__ counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new) return;
    counts[i] = (int32_t)er_paper<false>(n_exist, i, p, c, seed, yi, nullptr);
}

__global__ __launch_bounds__(256) void k_er_write(int64_t n_exist, int64_t n_new, double p, double c,
                                                  uint64_t seed, uint32_t yi,
                                                  const int64_t* __restrict__ new_off,
                                                  int64_t fwd_base, int32_t* __restrict__ fwd_off,
                                                  int32_t* __restrict__ fwd_adj,
                                                  int32_t* __restrict__ src_of) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new) return;
    const int64_t b = new_off[i];
    const int64_t e = er_paper<true>(n_exist, i, p, c, seed, yi, fwd_adj + fwd_base + b);
    for (int64_t k = 0; k < e; ++k) src_of[b + k] = (int32_t)(n_exist + i);
    fwd_off[n_exist + i + 1] = (int32_t)(fwd_base + new_off[i + 1]);
}

}  // namespace

void launch_er_count(hipStream_t s, int64_t n_exist, int64_t n_new, double p, uint64_t seed,
                     uint32_t year_index, int32_t* counts) {
    if (n_new <= 0) return;
    const double c = det_log1p(-p);
    int64_t nb = (n_new + 255) / 256;
    hipLaunchKernelGGL(k_er_count, dim3((uint32_t)nb), dim3(256), 0, s, n_exist, n_new, p, c, seed,
                       year_index, counts);
}

void launch_er_write(hipStream_t s, int64_t n_exist, int64_t n_new, double p, uint64_t seed,
                     uint32_t year_index, const int64_t* new_off, int64_t fwd_base,
                     int32_t* fwd_off, int32_t* fwd_adj, int32_t* src_of) {
    if (n_new <= 0) return;
    const double c = det_log1p(-p);
    int64_t nb = (n_new + 255) / 256;
    hipLaunchKernelGGL(k_er_write, dim3((uint32_t)nb), dim3(256), 0, s, n_exist, n_new, p, c, seed,
                       year_index, new_off, fwd_base, fwd_off, fwd_adj, src_of);
}

}  // namespace pg
