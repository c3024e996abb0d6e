/*
 * tomocam/array.h -- Array<T>, the move-only host container of the reference
 * API (include/array.h:40-261): row-major (n1, n2, n3), contiguous, with the
 * same constructors, indexing, arithmetic and factories.  Host arrays are the
 * interface only; every operator of the reconstruction path runs on the GPU.
 */
#ifndef ARRAY3__H
#define ARRAY3__H

#include <algorithm>
#include <complex>
#include <memory>
#include <random>
#include <span>
#include <stdexcept>
#include <type_traits>

#include "dtypes.h"
#include "slice.h"

namespace tomocam {

    template <typename T>
    class Array {
        dims_t dims_;
        dims_t pads_;
        size_t size_ = 0;
        std::unique_ptr<T[]> ptr_;

        template <typename F>
        Array &each(F f) {
            for (size_t i = 0; i < size_; ++i) ptr_[i] = f(ptr_[i]);
            return *this;
        }
        template <typename F>
        Array &each2(const Array &v, F f) {
            for (size_t i = 0; i < size_; ++i) ptr_[i] = f(ptr_[i], v.ptr_[i]);
            return *this;
        }

      public:
        Array() = default;
        Array(size_t a, size_t b, size_t c) : Array(dims_t(a, b, c)) {}
        explicit Array(dims_t d) : dims_(d), size_(d.size()), ptr_(std::make_unique<T[]>(d.size())) {}
        Array(const Array &) = delete;
        Array &operator=(const Array &) = delete;
        Array(Array &&o) noexcept : dims_(o.dims_), pads_(o.pads_), size_(o.size_), ptr_(std::move(o.ptr_)) {}
        Array &operator=(Array &&o) noexcept {
            if (this != &o) {
                dims_ = o.dims_;
                pads_ = o.pads_;
                size_ = o.size_;
                ptr_ = std::move(o.ptr_);
            }
            return *this;
        }

        [[nodiscard]] Array clone() const {
            Array r(dims_);
            std::copy(begin(), end(), r.begin());
            return r;
        }

        T *begin() { return ptr_.get(); }
        const T *begin() const { return ptr_.get(); }
        T *end() { return ptr_.get() + size_; }
        const T *end() const { return ptr_.get() + size_; }

        [[nodiscard]] dims_t dims() const { return dims_; }
        [[nodiscard]] size_t size() const { return size_; }
        [[nodiscard]] size_t nslices() const { return dims_.n1; }
        [[nodiscard]] size_t nrows() const { return dims_.n2; }
        [[nodiscard]] size_t ncols() const { return dims_.n3; }

        T &operator[](size_t i) { return ptr_[i]; }
        const T &operator[](size_t i) const { return ptr_[i]; }
        T &operator[](dims_t i) { return ptr_[dims_.flat_idx(i.n1, i.n2, i.n3)]; }
        const T &operator[](dims_t i) const { return ptr_[dims_.flat_idx(i.n1, i.n2, i.n3)]; }
        // zero outside the array (include/array.h:99-106)
        T at(size_t i, size_t j, size_t k) const {
            if (i >= dims_.n1 || j >= dims_.n2 || k >= dims_.n3) return T(0);
            return ptr_[dims_.flat_idx(i, j, k)];
        }

        Slice<T> slice(size_t b, size_t e) { return Slice<T>(begin() + b * dims_.n2 * dims_.n3, {e - b, dims_.n2, dims_.n3}); }
        const Slice<T> slice(size_t b, size_t e) const {
            return Slice<T>(const_cast<T *>(begin()) + b * dims_.n2 * dims_.n3, {e - b, dims_.n2, dims_.n3});
        }
        std::span<T> row(size_t i, size_t j) { return {begin() + (i * dims_.n2 + j) * dims_.n3, dims_.n3}; }
        const std::span<T> row(size_t i, size_t j) const {
            return {const_cast<T *>(begin()) + (i * dims_.n2 + j) * dims_.n3, dims_.n3};
        }

        void setPads(dims_t p) { pads_ = p; }
        [[nodiscard]] dims_t pads() const { return pads_; }

        // elementwise arithmetic (include/array.h:141-232)
        Array operator*(const Array &v) { return clone().each2(v, [](T a, T b) { return a * b; }); }
        Array &operator*=(T v) { return each([v](T a) { return a * v; }); }
        Array operator*(T v) { return clone() *= v; }
        Array &operator/=(T v) {
            if (std::abs(v) == 0) throw std::runtime_error("Division by zero in Array<T>::operator/=");
            return each([v](T a) { return a / v; });
        }
        Array &operator/=(const Array &v) {
            return each2(v, [](T a, T b) {
                if (b == T(0)) throw std::runtime_error("Division by zero in Array<T>::operator/=");
                return a / b;
            });
        }
        Array operator/(const Array &v) const { return clone() /= v; }
        Array operator/(T v) const { return clone() /= v; }
        Array &operator+=(T v) { return each([v](T a) { return a + v; }); }
        Array &operator+=(const Array &v) { return each2(v, [](T a, T b) { return a + b; }); }
        Array operator+(const Array &v) const { return clone() += v; }
        Array &operator-=(T v) { return each([v](T a) { return a - v; }); }
        Array &operator-=(const Array &v) { return each2(v, [](T a, T b) { return a - b; }); }
        Array operator-(const Array &v) const { return clone() -= v; }

        static Array zeros(const dims_t &d) {
            Array r(d);
            std::fill(r.begin(), r.end(), T(0));
            return r;
        }
        static Array ones(const dims_t &d) {
            Array r(d);
            std::fill(r.begin(), r.end(), T(1));
            return r;
        }
        // uniform [0, 1) from a randomly seeded Mersenne twister (include/array.h:245-258)
        static Array random(const dims_t &d) {
            static_assert(std::is_floating_point_v<T>, "Unsupported type for random");
            Array r(d);
            std::mt19937 gen(std::random_device{}());
            std::uniform_real_distribution<T> dis(T(0), T(1));
            for (auto &v : r) v = dis(gen);
            return r;
        }
    };

}  // namespace tomocam

#endif  // ARRAY3__H
