/*
 * tomocam/array_ops.h -- host reductions and conversions on Array<T>
 * (include/array_ops.h:39-93 of the reference).  Sums accumulate in double.
 * The solvers' own reductions run on the device (lam_dot, lam_max).
 */
#ifndef ARRAY_OPS_H
#define ARRAY_OPS_H

#include <algorithm>
#include <cmath>
#include <complex>

#include "array.h"

namespace tomocam::array {

    template <typename R>
    Array<std::complex<R>> to_complex(const Array<R> &a) {
        Array<std::complex<R>> b(a.dims());
        for (size_t i = 0; i < a.size(); ++i) b[i] = std::complex<R>(a[i]);
        return b;
    }

    template <typename R>
    Array<R> to_real(const Array<std::complex<R>> &a) {
        Array<R> b(a.dims());
        for (size_t i = 0; i < a.size(); ++i) b[i] = a[i].real();
        return b;
    }

    template <typename T>
    Array<T> abs(const Array<T> &a) {
        Array<T> b(a.dims());
        for (size_t i = 0; i < a.size(); ++i) b[i] = std::abs(a[i]);
        return b;
    }

    template <typename T>
    T max(const Array<T> &a) { return *std::max_element(a.begin(), a.end()); }

    template <typename T>
    T min(const Array<T> &a) { return *std::min_element(a.begin(), a.end()); }

    template <typename T>
    T dot(const Array<T> &a, const Array<T> &b) {
        double s = 0.0;
        for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
        return (T)s;
    }

    template <typename T>
    T norm2(const Array<T> &a) { return (T)std::sqrt((double)dot(a, a)); }

    template <typename T>
    T norm1(const Array<T> &a) {
        double s = 0.0;
        for (const T &v : a) s += std::abs((double)v);
        return (T)s;
    }

}  // namespace tomocam::array

#endif  // ARRAY_OPS_H
