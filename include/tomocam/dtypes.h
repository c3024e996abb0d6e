/*
 * tomocam/dtypes.h -- dims_t, the (n1, n2, n3) shape of a row-major volume
 * (same interface as include/dtypes.h:31-84 of the reference).
 *
 * The drop-in headers use the reference headers' include-guard names, so a
 * translation unit that also reaches a reference header of the same name
 * (through tiff.h, config.h, ...) gets one definition, whichever comes first.
 */
#ifndef DTYPES__H
#define DTYPES__H

#include <array>
#include <cstddef>
#include <cstdint>
#include <ostream>
#include <stdexcept>
#include <tuple>

namespace tomocam {

    struct dims_t {
        size_t n1 = 0, n2 = 0, n3 = 0;

        dims_t() = default;
        dims_t(size_t a, size_t b, size_t c) : n1(a), n2(b), n3(c) {}
        dims_t(const std::array<size_t, 3> &a) : n1(a[0]), n2(a[1]), n3(a[2]) {}

        [[nodiscard]] size_t size() const { return n1 * n2 * n3; }
        [[nodiscard]] size_t flat_idx(size_t i, size_t j, size_t k) const { return (i * n2 + j) * n3 + k; }
        [[nodiscard]] std::tuple<size_t, size_t, size_t> unravel_idx(size_t idx) const {
            return {idx / (n2 * n3), (idx / n3) % n2, idx % n3};
        }
        [[nodiscard]] size_t x() const { return n1; }
        [[nodiscard]] size_t y() const { return n2; }
        [[nodiscard]] size_t z() const { return n3; }
        void set_x(size_t v) { n1 = v; }
        void set_y(size_t v) { n2 = v; }
        void set_z(size_t v) { n3 = v; }

        dims_t operator+(const dims_t &o) const { return {n1 + o.n1, n2 + o.n2, n3 + o.n3}; }
        dims_t operator-(const dims_t &o) const { return {n1 - o.n1, n2 - o.n2, n3 - o.n3}; }
        dims_t operator*(int v) const { return {n1 * v, n2 * v, n3 * v}; }
        dims_t operator/(int v) const {
            if (v == 0) throw std::runtime_error("divide by zeros");
            return {n1 / v, n2 / v, n3 / v};
        }
        bool operator==(const dims_t &o) const { return n1 == o.n1 && n2 == o.n2 && n3 == o.n3; }
        bool operator!=(const dims_t &o) const { return !(*this == o); }
    };

    inline std::ostream &operator<<(std::ostream &os, dims_t d) {
        return os << "dims_t(" << d.n1 << ", " << d.n2 << ", " << d.n3 << ")";
    }

}  // namespace tomocam

#endif  // DTYPES__H
