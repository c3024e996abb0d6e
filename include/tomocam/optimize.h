/*
 * tomocam/optimize.h -- VecArray, Function and the solvers of
 * include/optimize.h:32-153 of the reference: split_bregman, cgsolver,
 * nagopt, lipschitz and qggmrf.  The solver loops run on the GPU
 * (lam_cgsolver, lam_split_bregman, lam_nagopt, lam_lipschitz) with FP64
 * scalars.  A host std::function operator is called through an adapter that
 * copies its argument to the host and its result back to the device each
 * time (correct, but slow): the MBIR drivers of tomocam.h keep the whole
 * loop on the device instead.
 */
#ifndef TOMOCAM_OPT_H
#define TOMOCAM_OPT_H

#include <array>
#include <cmath>
#include <exception>
#include <functional>
#include <random>

#include "array.h"
#include "array_ops.h"
#include "device.h"
#include "logger.h"
#include "projection.h"

namespace tomocam::opt {

    template <typename T>
    using VecArray = std::array<Array<T>, 3>;

    template <typename T>
    VecArray<T> operator+(const VecArray<T> &a, const VecArray<T> &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
    template <typename T>
    VecArray<T> &operator+=(VecArray<T> &a, const VecArray<T> &b) {
        for (int i = 0; i < 3; ++i) a[i] += b[i];
        return a;
    }
    template <typename T>
    VecArray<T> operator-(const VecArray<T> &a, const VecArray<T> &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
    template <typename T>
    VecArray<T> &operator-=(VecArray<T> &a, const VecArray<T> &b) {
        for (int i = 0; i < 3; ++i) a[i] -= b[i];
        return a;
    }
    template <typename T>
    VecArray<T> operator*(const VecArray<T> &a, T s) {
        return {a[0].clone() *= s, a[1].clone() *= s, a[2].clone() *= s};
    }

    template <typename T>
    using Function = std::function<VecArray<T>(const VecArray<T> &)>;

    template <typename T>
    using Residual = std::function<T(const VecArray<T> &)>;

    namespace detail {
        // Calls a host operator from the device solvers (lam_op_fn / lam_loss_fn).
        // Exceptions are held here and rethrown after the C call returns.
        template <typename T>
        struct HostOp {
            const Function<T> *op = nullptr;
            const Residual<T> *loss = nullptr;
            dims_t dims;
            std::exception_ptr err;

            VecArray<T> fetch(const float *const x[3]) const {
                VecArray<T> v{Array<T>(dims), Array<T>(dims), Array<T>(dims)};
                const size_t n = dims.size();
                std::vector<float> h(n);
                for (int i = 0; i < 3; ++i) {
                    device::check(lam_download(device::ctx(), h.data(), x[i], n * sizeof(float)));
                    device::check(lam_ctx_sync(device::ctx()));
                    for (size_t k = 0; k < n; ++k) v[i][k] = (T)h[k];
                }
                return v;
            }

            static lam_status apply(void *user, const float *const x[3], float *const y[3]) {
                auto *self = static_cast<HostOp *>(user);
                try {
                    VecArray<T> r = (*self->op)(self->fetch(x));
                    const size_t n = self->dims.size();
                    std::vector<float> h(n);
                    for (int i = 0; i < 3; ++i) {
                        if (r[i].size() != n) throw std::invalid_argument("operator returned a wrong-sized array");
                        for (size_t k = 0; k < n; ++k) h[k] = (float)r[i][k];
                        device::check(lam_upload(device::ctx(), y[i], h.data(), n * sizeof(float)));
                        device::check(lam_ctx_sync(device::ctx()));
                    }
                    return LAM_OK;
                } catch (...) {
                    self->err = std::current_exception();
                    return LAM_ERR_STATE;
                }
            }

            static lam_status value(void *user, const float *const x[3], double *out) {
                auto *self = static_cast<HostOp *>(user);
                try {
                    *out = (double)(*self->loss)(self->fetch(x));
                    return LAM_OK;
                } catch (...) {
                    self->err = std::current_exception();
                    return LAM_ERR_STATE;
                }
            }

            void finish(lam_status st) {
                if (err) std::rethrow_exception(err);
                device::check(st);
            }
        };

        template <typename T>
        dims_t common_dims(const VecArray<T> &v) {
            const dims_t d = v[0].dims();
            for (int i = 1; i < 3; ++i)
                if (v[i].dims() != d) throw std::invalid_argument("vector components differ in shape");
            return d;
        }
    }  // namespace detail

    // src/bregman.cpp:38-134
    template <typename T>
    VecArray<T> split_bregman(const Function<T> &A, const VecArray<T> &yT, const VecArray<T> &x0, T lambda, T mu,
                              size_t outer_max, size_t inner_max, T tol, T xtol) {
        if (x0.size() != yT.size() || x0[0].dims() != yT[0].dims())
            throw std::invalid_argument(
                "yT is not the raw data, but backprojected data, so it should have the same dimensions as x0");
        const dims_t d = detail::common_dims(x0);
        int64_t d64[3];
        tomocam::detail::dims64(d, d64);
        device::Vol3 dy = tomocam::detail::upload3(yT), dx0 = tomocam::detail::upload3(x0), dx(d.size());
        detail::HostOp<T> op{&A, nullptr, d, nullptr};
        op.finish(lam_split_bregman(device::ctx(), d64, &detail::HostOp<T>::apply, &op, dy.cptrs(), dx0.cptrs(),
                                    (double)lambda, (double)mu, (int64_t)outer_max, (int64_t)inner_max, (double)tol,
                                    (double)xtol, dx.ptrs()));
        return tomocam::detail::download3<T>(dx, d);
    }

    // src/conjgrad.cpp:36-117 (lambda is unused there as well)
    template <typename T>
    VecArray<T> cgsolver(const Function<T> &A, const VecArray<T> &b, const VecArray<T> &x0, size_t max_iters, T tol,
                         T lambda) {
        (void)lambda;
        const dims_t d = detail::common_dims(x0);
        if (detail::common_dims(b) != d) throw std::invalid_argument("b and x0 differ in shape");
        int64_t d64[3];
        tomocam::detail::dims64(d, d64);
        device::Vol3 db = tomocam::detail::upload3(b), dx0 = tomocam::detail::upload3(x0), dx(d.size());
        detail::HostOp<T> op{&A, nullptr, d, nullptr};
        int64_t iters = 0;
        op.finish(lam_cgsolver(device::ctx(), d64, &detail::HostOp<T>::apply, &op, db.cptrs(), dx0.cptrs(),
                               (int64_t)max_iters, (double)tol, dx.ptrs(), &iters));
        return tomocam::detail::download3<T>(dx, d);
    }

    // src/nagopt.cpp:115-209; x is updated in place and returned
    template <typename T>
    std::array<Array<T>, 3> nagopt(const std::function<std::array<Array<T>, 3>(const std::array<Array<T>, 3> &)> &grad,
                                   const std::function<T(const std::array<Array<T>, 3> &)> &loss,
                                   std::array<Array<T>, 3> &x, size_t max_iters, T lipschitz, T tol, T xtol,
                                   size_t max_inner_iters = 20, Logger *logger = nullptr) {
        (void)logger;
        const dims_t d = detail::common_dims(x);
        int64_t d64[3];
        tomocam::detail::dims64(d, d64);
        device::Vol3 dx = tomocam::detail::upload3(x);
        detail::HostOp<T> op{&grad, &loss, d, nullptr};
        op.finish(lam_nagopt(device::ctx(), d64, &detail::HostOp<T>::apply, &detail::HostOp<T>::value, &op,
                             dx.ptrs(), (int64_t)max_iters, (double)lipschitz, (double)tol, (double)xtol,
                             (int64_t)max_inner_iters));
        x = tomocam::detail::download3<T>(dx, d);
        return {x[0].clone(), x[1].clone(), x[2].clone()};
    }

    // src/lipschitz.cpp:30-53: power iteration from a random start
    template <typename T>
    T lipschitz(const Function<T> &grad, const Array<T> &x0, size_t max_iters = 20, T tol = 1e-5) {
        const dims_t d = x0.dims();
        int64_t d64[3];
        tomocam::detail::dims64(d, d64);
        detail::HostOp<T> op{&grad, nullptr, d, nullptr};
        double L = 0.0;
        const uint64_t seed = ((uint64_t)std::random_device{}() << 32) | std::random_device{}();
        op.finish(lam_lipschitz(device::ctx(), d64, &detail::HostOp<T>::apply, &op, (int64_t)max_iters, (double)tol,
                                seed, &L));
        return (T)L;
    }

    // src/qggmrf.cpp:59-95: g += gradient of the qGGMRF prior at x
    template <typename T>
    void qggmrf(Array<T> &g, const Array<T> &x, T sigma, T p) {
        if (g.dims() != x.dims()) throw std::invalid_argument("g and x differ in shape");
        int64_t d64[3];
        tomocam::detail::dims64(x.dims(), d64);
        auto dg = device::to_device(g.begin(), g.size());
        auto dx = device::to_device(x.begin(), x.size());
        device::check(lam_qggmrf(device::ctx(), d64, dg.get(), dx.get(), (float)sigma, (float)p));
        dg.download(g.begin());
    }

}  // namespace tomocam::opt

#endif  // TOMOCAM_OPT_H
