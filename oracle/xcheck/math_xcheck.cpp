// math_xcheck.cpp -- TEST INFRASTRUCTURE ONLY: the library's deterministic
// math header (paragraph_amd/csrc/pg_math.h, compiled here for the host)
// against the oracle's restatement (or_expf, or_logf, or_log).
//
//   math_xcheck expf [S]  every S-th of the 2^32 float bit patterns (S = 1:
//                         all of them)
//   math_xcheck logf [S]  likewise
//   math_xcheck log N     N doubles: random bit patterns, then positive
//                         normals for det_log_pos (which needs them)
// Prints "ok <count>" when every result has the oracle's bits, else the
// first mismatch, and exits non-zero.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../paragraph_amd/csrc/pg_math.h"
extern "C" {
#include "../oracle.h"
}

static uint32_t fbits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static uint64_t dbits(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }

static int sweep_f(bool is_exp, int64_t stride) {
    long long bad = 0;
    uint32_t first = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static, 1 << 20)
    for (int64_t i = 0; i < (int64_t(1) << 32) / stride; ++i) {
        const uint32_t u = (uint32_t)(i * stride);
        float x;
        std::memcpy(&x, &u, 4);
        const float a = is_exp ? pg::det_expf(x) : pg::det_logf(x);
        const float b = is_exp ? or_expf(x) : or_logf(x);
        if (fbits(a) != fbits(b) && !(a != a && b != b)) {
#pragma omp critical
            if (!bad) first = u;
            ++bad;
        }
    }
    if (bad) {
        std::printf("mismatch: %lld inputs, first 0x%08x\n", bad, first);
        return 1;
    }
    std::printf("ok %lld\n", (long long)((int64_t(1) << 32) / stride));
    return 0;
}

static int sweep_log(int64_t n) {
    uint64_t s = 0x9e3779b97f4a7c15ull;
    auto next = [&]() {  // splitmix64
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    for (int64_t i = 0; i < n; ++i) {
        uint64_t u = next();
        double x;
        std::memcpy(&x, &u, 8);
        const double a = pg::det_log(x), b = or_log(x);
        if (dbits(a) != dbits(b) && !(a != a && b != b)) {
            std::printf("det_log mismatch at 0x%016llx\n", (unsigned long long)u);
            return 1;
        }
        // a positive normal: exponent field in [1, 2046]
        u = (u & 0x800fffffffffffffull & ~0x8000000000000000ull) |
            ((uint64_t)(1 + next() % 2046) << 52);
        std::memcpy(&x, &u, 8);
        const double c = pg::det_log_pos(x), d = or_log(x);
        if (dbits(c) != dbits(d)) {
            std::printf("det_log_pos mismatch at 0x%016llx\n", (unsigned long long)u);
            return 1;
        }
    }
    std::printf("ok %lld\n", (long long)n);
    return 0;
}

int main(int argc, char** argv) {
    const int64_t stride = argc >= 3 ? std::atoll(argv[2]) : 1;
    if (stride < 1) return 2;
    if (argc >= 2 && !std::strcmp(argv[1], "expf")) return sweep_f(true, stride);
    if (argc >= 2 && !std::strcmp(argv[1], "logf")) return sweep_f(false, stride);
    if (argc >= 3 && !std::strcmp(argv[1], "log")) return sweep_log(std::atoll(argv[2]));
    std::fprintf(stderr, "usage: math_xcheck expf [S] | logf [S] | log N\n");
    return 2;
}
