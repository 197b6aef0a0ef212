// rocrand_philox.cpp -- TEST INFRASTRUCTURE ONLY: an independent Philox4x32-10
// implementation (rocRAND's engine, header-only, compiled for the host) that
// the tests compare the oracle's or_philox4x32_10 against.
//
// Reads lines "c0 c1 c2 c3 k0 k1" (counter words, key words) and prints the
// four output words.  rocRAND's engine is addressed by (seed, subsequence,
// offset): key = seed, counter words 2,3 = subsequence, counter words 0,1 =
// offset / 4, so c1 must stay below 2^30 (offset = 4 * (c0 | c1 << 32)).
#include <cstdio>

#include <rocrand/rocrand_philox4x32_10.h>

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        if (c1 >= (1u << 30)) {
            std::fprintf(stderr, "counter word 1 must be below 2^30\n");
            return 2;
        }
        const unsigned long long seed = (unsigned long long)k0 | ((unsigned long long)k1 << 32);
        const unsigned long long sub = (unsigned long long)c2 | ((unsigned long long)c3 << 32);
        const unsigned long long off = 4ull * ((unsigned long long)c0 | ((unsigned long long)c1 << 32));
        rocrand_device::philox4x32_10_engine e(seed, sub, off);
        const uint4 r = e.next4();
        std::printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
