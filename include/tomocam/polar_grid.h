/*
 * tomocam/polar_grid.h -- PolarGrid<T> (include/polar_grid.h:33-110 of the
 * reference).  Construction builds the device grid (lam_grid_polar): points
 * are generated on the fly by the kernels, never stored.  The public x, y, z
 * fields keep the reference's shape (P, R, C) and values but are materialised
 * on the host only when first read.
 */
#ifndef POLAR_GRID__H
#define POLAR_GRID__H

#include <memory>
#include <mutex>
#include <vector>

#include "array.h"
#include "device.h"

namespace tomocam {

    namespace detail {
        template <typename T>
        struct GridState {
            lam_grid *g = nullptr;
            double gamma = 0.0;
            dims_t dims;
            std::once_flag once;
            Array<T> xyz[3];

            ~GridState() {
                if (g) lam_grid_destroy(g);
            }
            const Array<T> &coord(int i) {
                std::call_once(once, [this] {
                    const size_t n = dims.size();
                    std::vector<float> h[3] = {std::vector<float>(n), std::vector<float>(n), std::vector<float>(n)};
                    device::check(lam_grid_coords_host(g, h[0].data(), h[1].data(), h[2].data()));
                    for (int a = 0; a < 3; ++a) {
                        xyz[a] = Array<T>(dims);
                        for (size_t k = 0; k < n; ++k) xyz[a][k] = (T)h[a][k];
                    }
                });
                return xyz[i];
            }
        };
    }  // namespace detail

    // Read-only view of one coordinate array of a PolarGrid.
    template <typename T>
    class GridCoord {
        std::shared_ptr<detail::GridState<T>> s_;
        int i_ = 0;

      public:
        GridCoord() = default;
        GridCoord(std::shared_ptr<detail::GridState<T>> s, int i) : s_(std::move(s)), i_(i) {}
        const Array<T> &get() const { return s_->coord(i_); }
        operator const Array<T> &() const { return get(); }
        dims_t dims() const { return s_ ? s_->dims : dims_t(); }
        size_t size() const { return dims().size(); }
        size_t nslices() const { return dims().n1; }
        size_t nrows() const { return dims().n2; }
        size_t ncols() const { return dims().n3; }
        const T *begin() const { return get().begin(); }
        const T *end() const { return get().end(); }
        const T &operator[](size_t k) const { return get()[k]; }
        const T &operator[](dims_t k) const { return get()[k]; }
    };

    template <typename T>
    struct PolarGrid {
        size_t npts = 0;
        std::vector<T> theta;
        GridCoord<T> x, y, z;

        PolarGrid() = default;

        // include/polar_grid.h:44-75: angles theta_i, detector (nrows, ncols), in-plane angle gamma
        PolarGrid(const std::vector<T> &angles, size_t nrows, size_t ncols, T gamma) : theta(angles) {
            auto s = std::make_shared<detail::GridState<T>>();
            s->gamma = (double)gamma;
            s->dims = dims_t(angles.size(), nrows, ncols);
            std::vector<double> th(angles.begin(), angles.end());
            device::check(lam_grid_polar(device::ctx(), th.data(), (int64_t)th.size(), (int64_t)nrows,
                                         (int64_t)ncols, (double)gamma, &s->g));
            npts = s->dims.size();
            x = GridCoord<T>(s, 0);
            y = GridCoord<T>(s, 1);
            z = GridCoord<T>(s, 2);
            state_ = std::move(s);
        }

        PolarGrid(const PolarGrid &) = delete;
        PolarGrid &operator=(const PolarGrid &) = delete;
        PolarGrid(PolarGrid &&) noexcept = default;
        PolarGrid &operator=(PolarGrid &&) noexcept = default;

        [[nodiscard]] dims_t dims() const { return x.dims(); }
        [[nodiscard]] size_t size() const { return x.size(); }
        [[nodiscard]] const std::vector<T> &angles() const { return theta; }
        [[nodiscard]] T angle(size_t i) const { return theta[i]; }
        [[nodiscard]] size_t nprojs() const { return theta.size(); }

        // device handle and the gamma the grid was built with
        lam_grid *handle() const {
            if (!state_) throw std::runtime_error("PolarGrid used before construction");
            return state_->g;
        }
        double gamma() const { return state_ ? state_->gamma : 0.0; }

      private:
        std::shared_ptr<detail::GridState<T>> state_;
    };

    // The device projectors take the beam direction from the grid's own gamma;
    // a different gamma at the call is rejected rather than silently ignored.
    template <typename T>
    void require_grid_gamma(const PolarGrid<T> &pg, T gamma) {
        if ((double)gamma != pg.gamma())
            throw std::invalid_argument("gamma differs from the gamma the PolarGrid was built with");
    }

}  // namespace tomocam

#endif  // POLAR_GRID__H
