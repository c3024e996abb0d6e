// Division of n in [0, 2^31) by a divisor d in [1, 2^31) fixed at launch time:
// one multiply-high, one add and one shift instead of the ~30-instruction
// integer division sequence.  Round-up magic number (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994):
//   s = ceil(log2 d),  m = floor(2^32 (2^s - d) / d) + 1,
//   n / d = (mulhi(n, m) + n) >> s.
// With e = (m + 2^32) d - 2^(32+s) in (0, d] and d <= 2^s the quotient is
// exact for every n < 2^32; n < 2^31 keeps the 32-bit add from overflowing.
// Needs no HIP header: a host test defines __host__ and __device__ away and
// checks it exhaustively.
#pragma once

namespace lam {

struct FastDiv {
    unsigned d = 1, m = 1, s = 0;

    FastDiv() = default;
    explicit FastDiv(unsigned divisor) : d(divisor) {
        s = 0;
        while ((1ull << s) < (unsigned long long)d) ++s;
        m = (unsigned)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
    }
    __host__ __device__ unsigned div(unsigned n) const {
        const unsigned hi = (unsigned)(((unsigned long long)n * m) >> 32);
        return (hi + n) >> s;
    }
    __host__ __device__ unsigned mod(unsigned n) const { return n - div(n) * d; }
};

}  // namespace lam
