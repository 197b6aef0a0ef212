// merge_host.cpp -- runs the device source of the backward-CSR merge
// (paragraph_amd/csrc/k_merge.inc, included verbatim) on the host and checks
// its rows against the oracle's or_csr_append.  TEST INFRASTRUCTURE ONLY.
//
// Every block's 256 threads are host threads: __syncthreads is a barrier of
// the block's threads, __shfl_up an exchange through a per-wavefront array
// between two barriers of the wavefront's 64 threads, the LDS atomicMax a
// compare-and-swap loop, __shared__ variables function statics (blocks run
// one after another).  Graphs with hubs (long backward runs) and long runs
// of papers without citers (blocks spanning more rows than positions) are
// merged with 1024, 2048 and 4096 positions per block, with and without the
// partition pass.  Exit status 0 when every variant equals the oracle.
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <thread>
#include <vector>

extern "C" {
#include "../oracle.h"
}

namespace emu {
struct Dim { unsigned x = 0, y = 0, z = 0; };
thread_local Dim threadIdx, blockIdx;
std::barrier<>* block_bar = nullptr;
std::barrier<>* wave_bar[4] = {};
long long slot[4][64];
}  // namespace emu

using emu::blockIdx;
using emu::threadIdx;
const emu::Dim blockDim{256, 1, 1};
using std::max;
using std::min;

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __restrict__ __restrict
#define __shared__ static

inline void __syncthreads() { emu::block_bar->arrive_and_wait(); }

template <typename T>
T __shfl_up(T v, int d, int) {
    const int w = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63;
    emu::slot[w][l] = (long long)v;
    emu::wave_bar[w]->arrive_and_wait();
    const T r = l >= d ? (T)emu::slot[w][l - d] : v;
    emu::wave_bar[w]->arrive_and_wait();
    return r;
}

inline int atomicMax(int* p, int v) {
    std::atomic_ref<int> a(*p);
    int cur = a.load();
    while (cur < v && !a.compare_exchange_weak(cur, v)) {
    }
    return cur;
}

#include "../../paragraph_amd/csrc/k_merge.inc"

namespace {

struct Case {
    int64_t n_old, n_new;
    std::vector<int32_t> fo, fa, bo, ba;   // old CSR
    std::vector<int32_t> noff, ntgt;       // new rows (citing new paper i)
};

// Old graph: papers cite older ones, a few hubs take most citations; new
// papers cite hubs and random older papers; papers past `quiet` are never
// cited (long runs of empty backward rows).
Case make_case(uint64_t seed, int64_t n_old, int64_t n_new, int hubs, int64_t quiet) {
    std::mt19937_64 rng(seed);
    Case c;
    c.n_old = n_old;
    c.n_new = n_new;
    std::vector<std::vector<int32_t>> rows(n_old);
    auto pick = [&](int64_t below) -> int32_t {
        const int64_t lim = std::min(below, quiet);
        if (lim <= 0) return -1;
        if (rng() % 3 == 0) return (int32_t)(rng() % std::min<int64_t>(hubs, lim));
        return (int32_t)(rng() % lim);
    };
    for (int64_t i = 1; i < n_old; ++i) {
        const int k = (int)(rng() % 12);
        for (int j = 0; j < k; ++j) {
            const int32_t t = pick(i);
            if (t >= 0) rows[i].push_back(t);
        }
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
    }
    c.fo.assign(n_old + 1, 0);
    for (int64_t i = 0; i < n_old; ++i) c.fo[i + 1] = c.fo[i] + (int32_t)rows[i].size();
    for (auto& r : rows) c.fa.insert(c.fa.end(), r.begin(), r.end());
    std::vector<std::vector<int32_t>> brow(n_old);
    for (int64_t i = 0; i < n_old; ++i)
        for (int32_t t : rows[i]) brow[t].push_back((int32_t)i);
    c.bo.assign(n_old + 1, 0);
    for (int64_t v = 0; v < n_old; ++v) c.bo[v + 1] = c.bo[v] + (int32_t)brow[v].size();
    for (auto& r : brow) c.ba.insert(c.ba.end(), r.begin(), r.end());
    c.noff.assign(n_new + 1, 0);
    for (int64_t i = 0; i < n_new; ++i) {
        std::vector<int32_t> r;
        const int k = (int)(rng() % 30);
        for (int j = 0; j < k; ++j) {
            const int32_t t = pick(n_old);
            if (t >= 0) r.push_back(t);
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        c.ntgt.insert(c.ntgt.end(), r.begin(), r.end());
        c.noff[i + 1] = (int32_t)c.ntgt.size();
    }
    return c;
}

template <int kOut>
bool run_merge(const Case& c, bool use_part, const std::vector<int32_t>& want) {
    const int64_t N = c.n_old + c.n_new, E = (int64_t)c.ntgt.size();
    // the year's (target, citer) pairs sorted by target, citers ascending
    std::vector<std::pair<uint32_t, uint32_t>> kv;
    for (int64_t i = 0; i < c.n_new; ++i)
        for (int32_t e = c.noff[i]; e < c.noff[i + 1]; ++e)
            kv.emplace_back((uint32_t)c.ntgt[e], (uint32_t)(c.n_old + i));
    std::stable_sort(kv.begin(), kv.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<uint32_t> src(kv.size());
    for (size_t i = 0; i < kv.size(); ++i) src[i] = kv[i].second;
    // new backward offsets: old offset + new citations of earlier papers
    std::vector<int32_t> newoff(N + 1);
    size_t p = 0;
    for (int64_t v = 0; v <= N; ++v) {
        while (p < kv.size() && kv[p].first < (uint32_t)v) ++p;
        newoff[v] = c.bo[std::min(v, c.n_old)] + (int32_t)p;
    }
    const int64_t e_new = c.bo[c.n_old] + E;
    std::vector<int32_t> out(e_new, -1);
    const int64_t nb = (e_new + kOut - 1) / kOut;
    std::vector<int64_t> part(2 * nb);
    if (use_part)
        for (int64_t b = 0; b < nb; ++b) {
            threadIdx.x = (unsigned)(b % 256);
            blockIdx.x = (unsigned)(b / 256);
            k_merge_part<kOut>(N, newoff.data(), e_new, nb, part.data());
        }
    std::barrier<> bb(256);
    std::barrier<> w0(64), w1(64), w2(64), w3(64);
    emu::block_bar = &bb;
    emu::wave_bar[0] = &w0; emu::wave_bar[1] = &w1; emu::wave_bar[2] = &w2; emu::wave_bar[3] = &w3;
    for (int64_t b = 0; b < nb; ++b) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < 256; ++t)
            th.emplace_back([&, t] {
                threadIdx.x = t;
                blockIdx.x = (unsigned)b;
                k_bwd_merge<kOut>(c.n_old, N, c.bo.data(), c.ba.data(), newoff.data(), src.data(),
                                  out.data(), e_new, use_part ? part.data() : nullptr);
            });
        for (auto& x : th) x.join();
    }
    const bool ok = out == want;
    std::printf("  positions per block %d, partition %s: %s\n", kOut, use_part ? "yes" : "no",
                ok ? "equal to the oracle" : "DIFFERENT");
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    bool ok = true;
    struct Spec { uint64_t seed; int64_t n_old, n_new; int hubs; int64_t quiet; };
    // `small`: graphs of a few blocks per variant (the CPU test suite)
    const bool small = argc > 1 && std::string(argv[1]) == "small";
    const Spec big[] = {{1, 6000, 2500, 8, 6000}, {2, 20000, 4000, 3, 2500}, {3, 3000, 9000, 1, 3000}};
    const Spec little[] = {{1, 1500, 600, 4, 1500}, {2, 5000, 800, 2, 600}, {3, 700, 2000, 1, 700}};
    for (const Spec& s : small ? little : big) {
        const Case c = make_case(s.seed, s.n_old, s.n_new, s.hubs, s.quiet);
        const int64_t N = c.n_old + c.n_new, E = (int64_t)c.ntgt.size();
        std::vector<int32_t> fo(N + 1), fa(c.fa.size() + E), bo(N + 1), ba(c.ba.size() + E), deg(N);
        or_csr g{c.n_old, c.fo.data(), c.fa.data(), c.bo.data(), c.ba.data()};
        if (or_csr_append(&g, c.n_new, c.noff.data(), c.ntgt.data(), fo.data(), fa.data(), bo.data(),
                          ba.data(), deg.data())) {
            std::printf("oracle append failed\n");
            return 2;
        }
        std::printf("graph %llu: %lld old papers, %lld new, %zu old and %lld new citations\n",
                    (unsigned long long)s.seed, (long long)c.n_old, (long long)c.n_new,
                    c.ba.size(), (long long)E);
        for (bool up : {false, true}) {
            ok = run_merge<1024>(c, up, ba) && ok;
            ok = run_merge<2048>(c, up, ba) && ok;
            ok = run_merge<4096>(c, up, ba) && ok;
        }
    }
    std::printf(ok ? "all variants equal the oracle\n" : "MISMATCH\n");
    return ok ? 0 : 1;
}
