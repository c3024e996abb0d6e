/*
 * tomocam/projection.h -- beam_dir_vector, forward and adjoint
 * (include/projection.h:36-67, src/projection.cpp:41-126 of the reference),
 * computed on the GPU: NUFFT type 2/1 of the three components, the centred
 * 2-D detector transform and the 1/(R*C) scale, as lam_forward/lam_adjoint.
 */
#ifndef PROJECTION__H
#define PROJECTION__H

#include <array>
#include <cmath>

#include "array.h"
#include "device.h"
#include "dtypes.h"
#include "polar_grid.h"

#ifndef TOMOCAM_NUFFT_EPS
#define TOMOCAM_NUFFT_EPS 1e-5   // NUFFT tolerance of the device projectors
#endif

namespace tomocam {

    // (sin g sin a, cos a, cos g sin a)
    template <typename T>
    inline std::array<T, 3> beam_dir_vector(T gamma, T alpha) {
        return {std::sin(gamma) * std::sin(alpha), std::cos(alpha), std::cos(gamma) * std::sin(alpha)};
    }

    namespace detail {
        inline void dims64(const dims_t &d, int64_t out[3]) {
            out[0] = (int64_t)d.n1; out[1] = (int64_t)d.n2; out[2] = (int64_t)d.n3;
        }

        template <typename T>
        device::Vol3 upload3(const std::array<Array<T>, 3> &v) {
            const dims_t d = v[0].dims();
            for (int i = 1; i < 3; ++i)
                if (v[i].dims() != d) throw std::invalid_argument("vector components differ in shape");
            device::Vol3 out(d.size());
            for (int i = 0; i < 3; ++i) out.c[i].upload(v[i].begin());
            return out;
        }

        template <typename T>
        std::array<Array<T>, 3> download3(const device::Vol3 &v, const dims_t &d) {
            std::array<Array<T>, 3> out{Array<T>(d), Array<T>(d), Array<T>(d)};
            for (int i = 0; i < 3; ++i) v.c[i].download(out[i].begin());
            return out;
        }
    }  // namespace detail

    template <typename T>
    Array<T> forward(const std::array<Array<T>, 3> &m, const PolarGrid<T> &pg, const T &gamma) {
        require_grid_gamma(pg, gamma);
        int64_t d[3];
        detail::dims64(m[0].dims(), d);
        device::Vol3 dm = detail::upload3(m);
        device::Buffer<float> p(pg.size());
        device::check(lam_forward(device::ctx(), pg.handle(), d, TOMOCAM_NUFFT_EPS, dm.c[0].get(), dm.c[1].get(),
                                  dm.c[2].get(), p.get()));
        Array<T> out(pg.dims());
        p.download(out.begin());
        return out;
    }

    template <typename T>
    std::array<Array<T>, 3> adjoint(const Array<T> &proj, const PolarGrid<T> &pg, const dims_t &recon_dims, T gamma) {
        require_grid_gamma(pg, gamma);
        if (proj.dims() != pg.dims()) throw std::invalid_argument("projection stack does not match the PolarGrid");
        int64_t d[3];
        detail::dims64(recon_dims, d);
        auto p = device::to_device(proj.begin(), proj.size());
        device::Vol3 m(recon_dims.size());
        device::check(lam_adjoint(device::ctx(), pg.handle(), d, TOMOCAM_NUFFT_EPS, p.get(), m.c[0].get(),
                                  m.c[1].get(), m.c[2].get()));
        return detail::download3<T>(m, recon_dims);
    }

}  // namespace tomocam

#endif  // PROJECTION__H
