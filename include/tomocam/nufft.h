/*
 * tomocam/nufft.h -- nufft::nufft3d1 / nufft3d2 (include/nufft.h:37-65 of the
 * reference): type 1 (iflag +1) and type 2 (iflag -1) 3-D NUFFTs between the
 * PolarGrid's points and a complex volume whose dims (n1, n2, n3) are the
 * modes (k3, k2, k1), FINUFFT mode order 0.  Computed on the GPU in complex64.
 */
#ifndef NUFFT__H
#define NUFFT__H

#include <complex>

#include "array.h"
#include "device.h"
#include "polar_grid.h"
#include "projection.h"

namespace tomocam::nufft {

    template <typename T>
    void nufft3d1(const Array<std::complex<T>> &cz, Array<std::complex<T>> &fz, const PolarGrid<T> &pg) {
        if (cz.size() != pg.size()) throw std::invalid_argument("cz does not match the PolarGrid");
        int64_t d[3];
        tomocam::detail::dims64(fz.dims(), d);
        auto c = device::to_device(cz.begin(), cz.size());
        device::Buffer<std::complex<float>> f(fz.size());
        device::check(lam_nufft3d1(device::ctx(), pg.handle(), d, TOMOCAM_NUFFT_EPS, c.get(), f.get()));
        f.download(fz.begin());
    }

    template <typename T>
    void nufft3d2(Array<std::complex<T>> &cz, const Array<std::complex<T>> &fz, const PolarGrid<T> &pg) {
        if (cz.size() != pg.size()) throw std::invalid_argument("cz does not match the PolarGrid");
        int64_t d[3];
        tomocam::detail::dims64(fz.dims(), d);
        auto f = device::to_device(fz.begin(), fz.size());
        device::Buffer<std::complex<float>> c(cz.size());
        device::check(lam_nufft3d2(device::ctx(), pg.handle(), d, TOMOCAM_NUFFT_EPS, c.get(), f.get()));
        c.download(cz.begin());
    }

}  // namespace tomocam::nufft

#endif  // NUFFT__H
