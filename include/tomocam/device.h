/*
 * tomocam/device.h -- the device side of the C++ drop-in layer.
 *
 * Every function of the tomocam:: headers in this directory runs on the GPU
 * through the C ABI of lamino.h.  This header owns the per-thread library
 * context, maps C status codes back to the exceptions the reference throws,
 * and moves host arrays to and from device memory (float32 on the device;
 * double host arrays are refused unless LAMINO_FP32_FOR_DOUBLE=1 accepts
 * FP32 arithmetic, and are then converted on the way).
 *
 * Link with -llamino_amd (lamino_amd/liblamino_amd.so).
 */
#ifndef TOMOCAM_DEVICE_H
#define TOMOCAM_DEVICE_H

#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../lamino.h"

namespace tomocam::device {

    // Throws the exception the reference throws for the same failure
    // (include/nufft.h:48,63, include/finufft_plan_cache.h:63-64,
    // include/fft.h:60, include/array.h:164, src/bregman.cpp:45).
    [[noreturn]] inline void raise(lam_status st) {
        const char *m = lam_last_error();
        const std::string msg = m ? m : "";
        switch (st) {
        case LAM_ERR_NOMEM: throw std::bad_alloc();
        case LAM_ERR_NUFFT: throw std::runtime_error("Error in finufft_execute");
        case LAM_ERR_TYPE: throw std::invalid_argument("Only FINUFFT type 1 and 2 are supported");
        case LAM_ERR_DIVZERO: throw std::runtime_error("Division by zero in Array<T>::operator/=");
        case LAM_ERR_ARG:
        case LAM_ERR_SHAPE: throw std::invalid_argument(msg);
        default: throw std::runtime_error(msg.empty() ? "lamino: device call failed" : msg);
        }
    }

    inline void check(lam_status st) {
        if (st != LAM_OK) raise(st);
    }

    // One library context per host thread, on the device named by LAMINO_DEVICE
    // (default 0).  The reference is not re-entrant (one global FINUFFT plan per
    // type); here each thread gets its own plans and stream.
    class Context {
        lam_ctx *h_ = nullptr;

      public:
        explicit Context(int device) { check(lam_ctx_create(device, &h_)); }
        ~Context() {
            if (h_) lam_ctx_destroy(h_);
        }
        Context(const Context &) = delete;
        Context &operator=(const Context &) = delete;
        lam_ctx *get() const { return h_; }
    };

    inline lam_ctx *ctx() {
        thread_local Context c([] {
            const char *d = std::getenv("LAMINO_DEVICE");
            return d ? std::atoi(d) : 0;
        }());
        return c.get();
    }

    // Device element type of a host element type: float for float/double,
    // complex<float> for complex<float>/complex<double>.
    template <typename T> struct dev_elem { using type = float; };
    template <typename R> struct dev_elem<std::complex<R>> { using type = std::complex<float>; };

    // The device path computes in FP32 with an ES kernel of width 6 (eps 1e-5),
    // while the reference's double instantiations run FINUFFT at 1e-14
    // (include/finufft_plan.h:42).  A double host array is therefore refused
    // unless the caller accepts FP32 arithmetic by setting
    // LAMINO_FP32_FOR_DOUBLE=1; the results then come back converted to double.
    inline void require_fp32_opt_in() {
        const char *v = std::getenv("LAMINO_FP32_FOR_DOUBLE");
        if (!v || std::strcmp(v, "1") != 0)
            throw std::invalid_argument(
                "lamino: double-precision arrays are computed in FP32 on the device (eps 1e-5, not the "
                "reference's 1e-14); set LAMINO_FP32_FOR_DOUBLE=1 to accept this");
    }

    // Move-only device buffer of n elements of D, freed on destruction.
    template <typename D>
    class Buffer {
        D *p_ = nullptr;
        size_t n_ = 0;

      public:
        Buffer() = default;
        explicit Buffer(size_t n) : n_(n) {
            void *p = nullptr;
            if (n) check(lam_malloc(ctx(), n * sizeof(D), &p));
            p_ = static_cast<D *>(p);
        }
        ~Buffer() {
            if (p_) lam_free(ctx(), p_);
        }
        Buffer(Buffer &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
        Buffer &operator=(Buffer &&o) noexcept {
            if (this != &o) {
                if (p_) lam_free(ctx(), p_);
                p_ = o.p_; n_ = o.n_;
                o.p_ = nullptr; o.n_ = 0;
            }
            return *this;
        }
        Buffer(const Buffer &) = delete;
        Buffer &operator=(const Buffer &) = delete;
        D *get() const { return p_; }
        size_t size() const { return n_; }

        // Host -> device, converting the element type when the host is double.
        template <typename H>
        void upload(const H *src) {
            if constexpr (sizeof(H) != sizeof(D)) require_fp32_opt_in();
            if constexpr (sizeof(H) == sizeof(D)) {
                check(lam_upload(ctx(), p_, src, n_ * sizeof(D)));
            } else {
                std::vector<D> tmp(n_);
                for (size_t i = 0; i < n_; ++i) tmp[i] = static_cast<D>(src[i]);
                check(lam_upload(ctx(), p_, tmp.data(), n_ * sizeof(D)));
            }
        }
        // Device -> host (synchronous), converting when the host is double.
        template <typename H>
        void download(H *dst) const {
            if constexpr (sizeof(H) == sizeof(D)) {
                check(lam_download(ctx(), dst, p_, n_ * sizeof(D)));
            } else {
                std::vector<D> tmp(n_);
                check(lam_download(ctx(), tmp.data(), p_, n_ * sizeof(D)));
                for (size_t i = 0; i < n_; ++i) dst[i] = static_cast<H>(tmp[i]);
            }
            check(lam_ctx_sync(ctx()));
        }
        void zero() { check(lam_memset(ctx(), p_, 0, n_ * sizeof(D))); }
    };

    template <typename H>
    Buffer<typename dev_elem<H>::type> to_device(const H *src, size_t n) {
        Buffer<typename dev_elem<H>::type> b(n);
        b.upload(src);
        return b;
    }

    // Three device volumes of n elements (a VecArray on the device).
    struct Vol3 {
        Buffer<float> c[3];
        explicit Vol3(size_t n) : c{Buffer<float>(n), Buffer<float>(n), Buffer<float>(n)} {}
        const float *const *cptrs() {
            cp_[0] = c[0].get(); cp_[1] = c[1].get(); cp_[2] = c[2].get();
            return cp_;
        }
        float *const *ptrs() {
            p_[0] = c[0].get(); p_[1] = c[1].get(); p_[2] = c[2].get();
            return p_;
        }

      private:
        const float *cp_[3] = {};
        float *p_[3] = {};
    };

}  // namespace tomocam::device

#endif  // TOMOCAM_DEVICE_H
