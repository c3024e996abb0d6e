// Host check of lamino_amd/csrc/fastdiv.h: the multiply-shift quotient and
// remainder equal the integer ones for every divisor 1..4096 over the first
// 20000 dividends and a sweep up to 2^31, and for large divisors over a
// strided sweep of [0, 2^31).  Exit status 0 when exact everywhere.
#define __host__
#define __device__
#include <cstdio>

#include "fastdiv.h"

int main() {
    unsigned long long bad = 0, checks = 0;
    for (unsigned d = 1; d <= 4096; ++d) {
        const lam::FastDiv f(d);
        for (unsigned n = 0; n < 20000; ++n, ++checks)
            if (f.div(n) != n / d || f.mod(n) != n % d) ++bad;
        for (unsigned long long n = (1ull << 31) - 20000; n < (1ull << 31); n += 3, ++checks)
            if (f.div((unsigned)n) != (unsigned)n / d) ++bad;
    }
    const unsigned big[] = {65537u, 1u << 20, 3u * 5u * 7u * 11u * 13u * 17u * 19u, 1234567891u, 1u << 30,
                            (1u << 31) - 1};
    for (unsigned d : big) {
        const lam::FastDiv f(d);
        for (unsigned long long n = 0; n < (1ull << 31); n += 65521, ++checks)
            if (f.div((unsigned)n) != (unsigned)n / d || f.mod((unsigned)n) != (unsigned)n % d) ++bad;
    }
    std::printf("fastdiv: %llu checks, %llu wrong\n", checks, bad);
    return bad != 0;
}
