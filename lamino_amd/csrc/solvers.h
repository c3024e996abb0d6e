// Device-resident solvers: PCG (src/conjgrad.cpp:36-117), split-Bregman
// (src/bregman.cpp:38-134), Nesterov with line search (src/nagopt.cpp:115-209)
// and the MBIR job drivers (src/mbir_cg.cpp, src/mbir_sb.cpp, src/mbir3.cpp).
// Scalars (dots, step sizes, convergence tests) are FP64.
#pragma once
#include <functional>

#include "ops.h"
#include "toeplitz.h"

namespace lam {

// y = A(x) on device volumes, ordered on the context stream
using OpFn = std::function<void(CVol3, Vol3)>;
using LossFn = std::function<double(CVol3)>;

struct CgStats {
    long long iters = 0;
    long long applies = 0;
};

// x <- cgsolver(A, b, x0) ; x may alias x0.
void cg_solve(Ctx &cx, const long long d[3], const OpFn &A, CVol3 b, CVol3 x0, Vol3 x, long long max_iters,
              double tol, CgStats *st);

// Ab(u) = A(u); the split-Bregman system adds mu*laplacian itself.  When
// A_has_lap is set, A already returns A(u) + mu*laplacian(u) (fused form).
void split_bregman(Ctx &cx, const long long d[3], const OpFn &A, bool A_has_lap, CVol3 yT, CVol3 x0, double lambda,
                   double mu, long long outer_max, long long inner_max, double tol, double xtol, Vol3 x,
                   CgStats *st, long long *outer_done);

void nagopt(Ctx &cx, const long long d[3], const OpFn &grad, const LossFn &loss, Vol3 x, long long max_iters,
            double lipschitz, double tol, double xtol, long long max_inner, long long *outer_done,
            long long *grad_calls);

double lipschitz(Ctx &cx, const long long d[3], const OpFn &A, long long max_iters, double tol, uint64_t seed);

// device buffer holding 3 volumes
struct Vol3Buf {
    DevBuf buf;
    long long n = 0;
    void alloc(long long nn) { n = nn; buf.alloc(sizeof(float) * 3 * (size_t)nn); }
    Vol3 v() const { return Vol3{{buf.as<float>(), buf.as<float>() + n, buf.as<float>() + 2 * n}}; }
    CVol3 cv() const { return CVol3{{buf.as<float>(), buf.as<float>() + n, buf.as<float>() + 2 * n}}; }
};

inline CVol3 cv(Vol3 v) { return CVol3{{v.c[0], v.c[1], v.c[2]}}; }

}  // namespace lam
