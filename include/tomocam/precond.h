/*
 * tomocam/precond.h -- opt::RampPreconditioner<T> (include/precond.h:33-87 of
 * the reference): per n1-slice 2-D ramp filter sqrt((i/n3)^2 + (j'/n2)^2) in
 * the Fourier domain, scaled by 1/(n2*n3), DC zeroed; computed on the GPU.
 */
#ifndef TOMOCAM_PRECOND_H
#define TOMOCAM_PRECOND_H

#include "array.h"
#include "device.h"
#include "projection.h"

namespace tomocam::opt {

    template <typename T>
    class RampPreconditioner {
        dims_t dims_;

      public:
        explicit RampPreconditioner(dims_t dims) : dims_(dims) {}

        Array<T> apply(const Array<T> &input) const {
            if (input.dims() != dims_) throw std::invalid_argument("input does not match the preconditioner's dims");
            int64_t d[3];
            tomocam::detail::dims64(dims_, d);
            auto in = device::to_device(input.begin(), input.size());
            device::Buffer<float> out(input.size());
            device::check(lam_precond(device::ctx(), d, in.get(), out.get()));
            Array<T> r(dims_);
            out.download(r.begin());
            return r;
        }
    };

}  // namespace tomocam::opt

#endif  // TOMOCAM_PRECOND_H
