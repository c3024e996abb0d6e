/*
 * tomocam/tomocam.h -- the job-level API of the reference (include/tomocam.h:
 * 50-154): sysmat, gradient, residual and the three reconstructions
 * MBIR1 (CG), MBIR2 (split-Bregman TV) and MBIR3 (NAG + qGGMRF).
 *
 * The MBIR drivers run the whole job on the GPU (lam_mbir): normalisation,
 * padding, backprojection, the exact Toeplitz normal operator, the solver
 * loop and the crop.  Only the projections go in and the three cropped
 * components come out.
 */
#ifndef TOMOCAM__H
#define TOMOCAM__H

#include <array>
#include <cmath>
#include <tuple>
#include <type_traits>
#include <vector>

#include "array.h"
#include "array_ops.h"
#include "bregman.h"
#include "config.h"
#include "device.h"
#include "dtypes.h"
#include "nufft.h"
#include "optimize.h"
#include "polar_grid.h"
#include "precond.h"
#include "projection.h"

namespace tomocam {

    template <typename T>
    concept Float = std::is_floating_point<T>::value;

    // src/gradient.cpp:33-89: A x = R^T R x, NUFFT form
    template <typename F>
    std::array<Array<F>, 3> sysmat(const std::array<Array<F>, 3> &x, const PolarGrid<F> &grid, F gamma) {
        require_grid_gamma(grid, gamma);
        const dims_t d = x[0].dims();
        int64_t d64[3];
        detail::dims64(d, d64);
        device::Vol3 dx = detail::upload3(x), dy(d.size());
        device::check(lam_sysmat(device::ctx(), grid.handle(), d64, TOMOCAM_NUFFT_EPS, dx.cptrs(), dy.ptrs()));
        return detail::download3<F>(dy, d);
    }

    // src/gradient.cpp:96-105: sysmat(m) - pT
    template <typename F>
    std::array<Array<F>, 3> gradient(const std::array<Array<F>, 3> &m, const std::array<Array<F>, 3> &pT,
                                     const PolarGrid<F> &grid, F gamma) {
        auto g = sysmat(m, grid, gamma);
        for (int i = 0; i < 3; ++i) g[i] -= pT[i];
        return g;
    }

    // src/gradient.cpp:115-127: sum_i sqrt(m_i.A m_i - 2 m_i.pT_i + pTp); the
    // dots run on the device in FP64.
    template <typename F>
    F residual(const std::array<Array<F>, 3> &m, const std::array<Array<F>, 3> &pT, const PolarGrid<F> &grid, F pTp,
               F gamma) {
        require_grid_gamma(grid, gamma);
        const dims_t d = m[0].dims();
        int64_t d64[3];
        detail::dims64(d, d64);
        device::Vol3 dm = detail::upload3(m), dp = detail::upload3(pT), dA(d.size());
        device::check(lam_sysmat(device::ctx(), grid.handle(), d64, TOMOCAM_NUFFT_EPS, dm.cptrs(), dA.ptrs()));
        F err = F(0);
        for (int i = 0; i < 3; ++i) {
            double xAx = 0.0, yx = 0.0;
            device::check(lam_dot(device::ctx(), (int64_t)d.size(), dm.c[i].get(), dA.c[i].get(), &xAx));
            device::check(lam_dot(device::ctx(), (int64_t)d.size(), dm.c[i].get(), dp.c[i].get(), &yx));
            err += std::sqrt((F)xAx - F(2) * (F)yx + pTp);
        }
        return err;
    }

    namespace detail {
        template <typename F>
        std::array<Array<F>, 3> mbir(int method, const std::vector<std::tuple<Array<F>, std::vector<F>, F>> &datasets,
                                     const dims_t &recon_dims, const ReconParams &params) {
            std::vector<std::vector<float>> proj(datasets.size());
            std::vector<std::vector<double>> alpha(datasets.size());
            std::vector<lam_dataset> ds(datasets.size());
            for (size_t g = 0; g < datasets.size(); ++g) {
                const auto &[p, angles, gamma] = datasets[g];
                if (p.nslices() != angles.size())
                    throw std::invalid_argument("number of angles does not match the projection stack");
                proj[g].assign(p.begin(), p.end());
                alpha[g].assign(angles.begin(), angles.end());
                ds[g] = lam_dataset{(int64_t)p.nslices(), (int64_t)p.nrows(), (int64_t)p.ncols(), proj[g].data(),
                                    alpha[g].data(), nullptr, (double)gamma};
            }
            lam_recon_params rp{};
            rp.method = method;
            rp.max_iters = (int64_t)params.maxIters;
            rp.inner_iters = (int64_t)params.innerIters;
            rp.lambda = params.lambda;
            rp.mu = params.mu;
            rp.sigma = params.sigma;
            rp.p = params.p;
            rp.tol = params.tol;
            rp.xtol = params.xtol;
            rp.pad_factor = params.PAD_FACTOR;
            rp.eps = TOMOCAM_NUFFT_EPS;
            int64_t rd[3];
            dims64(recon_dims, rd);
            std::vector<float> out(3 * recon_dims.size());
            device::check(lam_mbir(device::ctx(), ds.data(), (int)ds.size(), rd, &rp, out.data(), nullptr));
            std::array<Array<F>, 3> r{Array<F>(recon_dims), Array<F>(recon_dims), Array<F>(recon_dims)};
            for (int i = 0; i < 3; ++i)
                for (size_t k = 0; k < recon_dims.size(); ++k) r[i][k] = (F)out[i * recon_dims.size() + k];
            return r;
        }
    }  // namespace detail

    // src/mbir_cg.cpp:47-128
    template <typename F>
    std::array<Array<F>, 3> MBIR1(const std::vector<std::tuple<Array<F>, std::vector<F>, F>> &datasets,
                                  const dims_t &recon_dims, const ReconParams &params) {
        return detail::mbir<F>(1, datasets, recon_dims, params);
    }

    // src/mbir_sb.cpp:46-114
    template <typename F>
    std::array<Array<F>, 3> MBIR2(const std::vector<std::tuple<Array<F>, std::vector<F>, F>> &datasets,
                                  const dims_t &recon_dims, const ReconParams &params) {
        return detail::mbir<F>(2, datasets, recon_dims, params);
    }

    // src/mbir3.cpp:46-135
    template <typename F>
    std::array<Array<F>, 3> MBIR3(const std::vector<std::tuple<Array<F>, std::vector<F>, F>> &datasets,
                                  const dims_t &recon_dims, const ReconParams &params) {
        return detail::mbir<F>(3, datasets, recon_dims, params);
    }

}  // namespace tomocam

#endif  // TOMOCAM__H
