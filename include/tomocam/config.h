/*
 * tomocam/config.h -- reconstruction parameters and dataset tuple of the
 * reference (include/config.h:57-58, 143-160), without the TOML reader.
 */
#ifndef CONFIG_H
#define CONFIG_H

#include <array>
#include <cstddef>
#include <tuple>
#include <vector>

#include "array.h"

namespace tomocam {

    // (projections [P][R][C], angles per projection, gamma)
    template <typename T>
    using Dataset_t = std::tuple<Array<T>, std::vector<T>, T>;

    enum class Regularizer { qGGMRF, SPLIT_BREGMAN, UNCONSTRAINED };

    // Same fields and defaults as the reference's ReconParams.
    struct ReconParams {
        Regularizer regularizer = Regularizer::UNCONSTRAINED;
        std::array<size_t, 3> recon_dims = {0, 0, 0};
        size_t maxIters = 50;
        size_t innerIters = 3;
        float sigma = 1000.0f;
        float p = 1.2f;
        float lambda = 0.1f;
        float mu = 10.0f;
        float tol = 1e-5f;
        float xtol = 1e-5f;
        float PAD_FACTOR = 1.4142f;
    };

}  // namespace tomocam

#endif  // CONFIG_H
