// Shared helpers of the C ABI entry points: exception -> status mapping with a
// per-thread message, and argument checks.
#pragma once
#include <new>
#include <string>

#include "common.h"

namespace lam {

inline thread_local std::string g_last_error;

template <class F> lam_status guard(F &&f) {
    try {
        f();
        return LAM_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &e) {
        g_last_error = e.what();
        return LAM_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return LAM_ERR_ARG;
    }
}

inline void need(const void *p, const char *what) {
    if (!p) throw Error(LAM_ERR_ARG, std::string("null argument: ") + what);
}

inline void dims_of(const int64_t *d, long long out[3]) {
    need(d, "dims");
    for (int a = 0; a < 3; ++a) {
        if (d[a] < 1) throw Error(LAM_ERR_ARG, "dimensions must be positive");
        out[a] = d[a];
    }
}

}  // namespace lam
