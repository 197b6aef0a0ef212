// pg_noalloc.hip -- pg_bfs_noalloc: 2-hop neighbourhoods of roots held in
// caller device memory, written to caller device memory, with every scratch
// buffer of the launch in one file-scope device array.  No allocation, no
// host synchronisation, no memset: usable under stream capture and in loops
// that may not allocate.
//
// Scratch layout (int32 words of `scratch`):
//   [0, 64)                  counters: tier-list counts [0..1], root count
//                            (int64 at words 8-9)
//   [64, 64 + kNaSlotPad)    staging-slot busy flags (kNaSlots used)
//   next 2 * kNaMaxRoots     the wavefront tiers' hand-over lists
//   next kNaSlots * kBfsStageEntries   staging slots of the workgroup tier
// about 24 MiB per device, held until the process ends.
#include <mutex>
#include <string>

#include "../../include/paragraph.h"
#include "pg_ctx.h"
#include "pg_internal.h"

using namespace pg;

namespace {

constexpr int kNaSlots = 32;
constexpr int kNaSlotPad = 64;
constexpr int64_t kNaMaxRoots = int64_t(1) << 20;
constexpr int64_t kNaCounterWords = 64;
constexpr int64_t kNaListWords = 2 * kNaMaxRoots;
constexpr int64_t kNaSlotWords = (int64_t)kNaSlots * kBfsStageEntries;
constexpr int64_t kNaWords = kNaCounterWords + kNaSlotPad + kNaListWords + kNaSlotWords;
constexpr int kNaMaxDevices = 64;

}  // namespace

static __device__ __align__(256) int32_t scratch[kNaWords];

namespace {

std::mutex g_na_mu;
int32_t* g_na_ptr[kNaMaxDevices] = {};
hipError_t g_na_err[kNaMaxDevices] = {};

// Clears the tier-list counts and the busy flags, and stores the root count.
__global__ void k_na_init(int32_t* counters, int32_t* busy, int n_busy, int64_t count) {
    const int t = threadIdx.x;
    if (t < 2) counters[t] = 0;
    if (t == 2) *reinterpret_cast<int64_t*>(counters + 8) = count;
    for (int k = t; k < n_busy; k += blockDim.x) busy[k] = 0;
}

CsrView view_of(const pg_ctx* c) {
    return CsrView{c->fwd_off, c->fwd_adj, c->bwd_off[c->cur], c->bwd_adj[c->cur], c->n};
}

}  // namespace

// Device address of `scratch` on ctx's device (current), looked up once per
// device.  The array's name is only a host-side handle: kernels and copies
// must be given this address.
int noalloc_setup(pg_ctx* ctx) {
    const int dev = ctx->cfg.device;
    if (dev < 0 || dev >= kNaMaxDevices) return PG_OK;  // pg_bfs_noalloc refuses this device
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[dev]) return PG_OK;
    void* p = nullptr;
    hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p) {
        g_na_err[dev] = r != hipSuccess ? r : hipErrorInvalidSymbol;
        return PG_OK;  // reported by pg_bfs_noalloc; other entry points are unaffected
    }
    g_na_ptr[dev] = static_cast<int32_t*>(p);
    return PG_OK;
}

extern "C" int pg_bfs_noalloc(pg_ctx* ctx, int64_t count, const int32_t* roots_dev,
                              int32_t* out_dev, int32_t* out_len_dev, void* stream,
                              int32_t stream_given) {
    if (!ctx) return PG_EINVAL;
    if (!ctx->loaded) return fail(ctx, PG_ESTATE, "pg_bfs_noalloc: no graph loaded");
    if (count < 0 || (count > 0 && (!roots_dev || !out_dev || !out_len_dev)))
        return fail(ctx, PG_EINVAL, "pg_bfs_noalloc: bad arguments");
    if (count > kNaMaxRoots)
        return fail(ctx, PG_ECAP,
                    "pg_bfs_noalloc: count exceeds the 1048576 roots its static scratch holds "
                    "(split the call)");
    if (count == 0) return PG_OK;
    const int dev = ctx->cfg.device;
    int32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_na_mu);
        if (dev >= 0 && dev < kNaMaxDevices) base = g_na_ptr[dev];
        if (!base) {
            const hipError_t e = dev >= 0 && dev < kNaMaxDevices ? g_na_err[dev] : hipErrorInvalidDevice;
            return fail(ctx, PG_EHIP, std::string("pg_bfs_noalloc: no scratch address on this "
                                                  "device: ") + hipGetErrorString(e));
        }
    }
    PG_HIP(hipSetDevice(dev));
    hipStream_t s = stream_given ? static_cast<hipStream_t>(stream) : ctx->stream;
    int32_t* counters = base;
    int32_t* busy = base + kNaCounterWords;
    int32_t* lists = busy + kNaSlotPad;
    BfsStage pool;
    pool.buf = lists + kNaListWords;
    pool.busy = busy;
    pool.n_slots = kNaSlots;
    pool.cap_entries = kBfsStageEntries;
    hipLaunchKernelGGL(k_na_init, dim3(1), dim3(64), 0, s, counters, busy, kNaSlots, count);
    // The three tiers of the year step without weights: wavefront per root
    // (small, then medium), workgroup per remaining root; S and its length
    // per root, no CDF.  The hand-over lists hold count entries each.
    launch_bfs_root(s, view_of(ctx), nullptr, roots_dev,
                    reinterpret_cast<const int64_t*>(counters + 8), count, ctx->cfg.bfs_cap,
                    out_dev, nullptr, nullptr, out_len_dev, lists, counters, nullptr,
                    bfs_stage_view(pool, ctx->tune_bfs_mode, ctx->tune_small_max,
                                   ctx->tune_mid_max),
                    nullptr);
    PG_HIP(hipGetLastError());
    return PG_OK;
}
