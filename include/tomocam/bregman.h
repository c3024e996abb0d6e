/*
 * tomocam/bregman.h -- the finite-difference operators of the split-Bregman
 * solver (include/bregman.h:15-36, src/finitediff.cpp:7-71 of the
 * reference), on the GPU: central differences on interior voxels, zero on
 * every boundary face.
 */
#ifndef BREGMAN_H
#define BREGMAN_H

#include <array>

#include "array.h"
#include "device.h"
#include "projection.h"

namespace tomocam::opt {

    template <typename T>
    std::array<Array<T>, 3> grad_u(const Array<T> &u) {
        int64_t d[3];
        tomocam::detail::dims64(u.dims(), d);
        auto du = device::to_device(u.begin(), u.size());
        device::Vol3 g(u.size());
        device::check(lam_grad_u(device::ctx(), d, du.get(), g.c[0].get(), g.c[1].get(), g.c[2].get()));
        return tomocam::detail::download3<T>(g, u.dims());
    }

    template <typename T>
    Array<T> divergence(const std::array<Array<T>, 3> &v) {
        const dims_t dims = v[0].dims();
        int64_t d[3];
        tomocam::detail::dims64(dims, d);
        device::Vol3 dv = tomocam::detail::upload3(v);
        device::Buffer<float> out(dims.size());
        device::check(lam_divergence(device::ctx(), d, dv.c[0].get(), dv.c[1].get(), dv.c[2].get(), out.get()));
        Array<T> r(dims);
        out.download(r.begin());
        return r;
    }

    // divergence(grad_u(u)), fused on the device
    template <typename T>
    Array<T> laplacian(const Array<T> &u) {
        int64_t d[3];
        tomocam::detail::dims64(u.dims(), d);
        auto du = device::to_device(u.begin(), u.size());
        device::Buffer<float> out(u.size());
        device::check(lam_laplacian(device::ctx(), d, du.get(), out.get()));
        Array<T> r(u.dims());
        out.download(r.begin());
        return r;
    }

}  // namespace tomocam::opt

#endif  // BREGMAN_H
