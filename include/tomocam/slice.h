/*
 * tomocam/slice.h -- Slice<T>, a non-owning view of consecutive n1-slices of
 * an Array (include/slice.h:30-63 of the reference).
 */
#ifndef SLICE__H
#define SLICE__H

#include "dtypes.h"

namespace tomocam {

    template <typename T>
    class Slice {
        T *data_;
        dims_t dims_;

      public:
        Slice(T *data, dims_t d) : data_(data), dims_(d) {}
        dims_t dims() const { return dims_; }
        size_t size() const { return dims_.size(); }
        T *begin() { return data_; }
        T *end() { return data_ + size(); }
        const T *begin() const { return data_; }
        const T *end() const { return data_ + size(); }
        T &operator[](dims_t i) { return data_[dims_.flat_idx(i.n1, i.n2, i.n3)]; }
        const T &operator[](dims_t i) const { return data_[dims_.flat_idx(i.n1, i.n2, i.n3)]; }
    };

}  // namespace tomocam

#endif  // SLICE__H
