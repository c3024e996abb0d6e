// Fused elementwise, stencil and reduction kernels on [n1][n2][n3] volumes.
#pragma once
#include "common.h"
#include "fastdiv.h"

namespace lam {

struct Vol3 {
    float *c[3];
};
struct CVol3 {
    const float *c[3];
};

// ---- reductions (FP64, fixed launch shape => deterministic order)
// Term kinds: sum a*b, sum (a-b)^2, sum a^2
enum { RED_DOT = 0, RED_DIFF2 = 1, RED_SQ = 2 };
struct RedTerm {
    int kind;
    const float *a, *b;
};
constexpr int kMaxRedTerms = 12;
// Computes sums of up to 12 terms over n elements; results to host (blocking).
void reduce_terms(Ctx &cx, long long n, const RedTerm *terms, int nterms, double *out_host);
// Same, results left in a device double array (no sync).  partials (>= kRedPartials
// doubles) replaces the context's shared scratch, for work captured into graphs.
constexpr int kRedBlocks = 1024;
constexpr size_t kRedPartials = (size_t)kRedBlocks * kMaxRedTerms;
void reduce_terms_dev(Ctx &cx, long long n, const RedTerm *terms, int nterms, double *out_dev,
                      double *partials = nullptr);
float max_value(Ctx &cx, long long n, const float *a);
// sum a*b in FP64 like reduce_terms, with its partial sums in a static device
// array instead of context memory: no device allocation (n <= kNoAllocMaxElems)
constexpr long long kNoAllocMaxElems = 1ll << 40;
// non-finite values -> 0 (include/mask.h:22-32); out[j][k][i] = in[i][j][k]
void mask_infs_nans(Ctx &cx, long long n, float *a);
void transpose120_dev(Ctx &cx, const long long d[3], const float *in, float *out);
// out (pd, zeroed by the caller) = pad3d(transpose(in, {2, 0, 1})) for in [d0][d1][d2]
void transpose201_pad(Ctx &cx, const long long d[3], const float *in, const long long pd[3], float *out);
// crop2d SYMMETRIC of P planes R2 x C2 to R x C
void crop2(Ctx &cx, const float *in, long long P, int R2, int C2, float *out, int R, int C);
// demagnetising field of a 3-component volume (include/demag.h); work: 3 n complex
void demag(Ctx &cx, const long long d[3], const float *const m[3], float *const h[3], float2 *work);
double dot_noalloc(Ctx &cx, long long n, const float *a, const float *b);

// ---- stencils (src/finitediff.cpp:7-71, include/bregman.h:33-36)
void grad_u(Ctx &cx, const long long d[3], const float *u, float *g0, float *g1, float *g2);
void divergence(Ctx &cx, const long long d[3], const float *v0, const float *v1, const float *v2, float *out);
// out = laplacian(u)  or  y_c += mu * laplacian(x_c) for the 3 components
void laplacian(Ctx &cx, const long long d[3], const float *u, float *out);
void add_mu_laplacian3(Ctx &cx, const long long d[3], CVol3 x, Vol3 y, float mu);

// ---- qGGMRF prior gradient (src/qggmrf.cpp:59-95): g_c += prior(x_c), c < ncomp
void qggmrf(Ctx &cx, const long long d[3], float *g, const float *x, float sigma, float p);
// the same for components 0..ncomp-1 of x into those of g, in one launch
void qggmrf3(Ctx &cx, const long long d[3], Vol3 g, CVol3 x, int ncomp, float sigma, float p);

// ---- ramp preconditioner (include/precond.h:33-87), out = M in, on 1 or 2 volumes at once.
// work (>= n complex) replaces the context's shared workspace, for work captured into graphs.
void precond_apply(Ctx &cx, const long long d[3], const float *in0, const float *in1, float *out0, float *out1,
                   float2 *work = nullptr);

// ---- elementwise helpers
// y = a*x + b*y over n floats (3 comps, optional)
void axpby3(Ctx &cx, long long n, float a, CVol3 x, float b, Vol3 y);
void copy3(Ctx &cx, long long n, CVol3 x, Vol3 y);
void fill(Ctx &cx, long long n, float *y, float v);
void scale_to(Ctx &cx, long long n, const float *x, float s, float *y);

// (i, j, k) of a flat index, k fastest: multiply-shift splits when the volume has
// fewer than 2^31 cells (a 64-bit division costs ~40 instructions), else division.
struct Split3 {
    FastDiv f23, f3;
    int small = 0;
};
inline Split3 make_split3(const long long d[3]) {
    Split3 s;
    s.small = d[0] * d[1] * d[2] < (1ll << 31) ? 1 : 0;
    if (s.small) {
        s.f23 = FastDiv((unsigned)(d[1] * d[2]));
        s.f3 = FastDiv((unsigned)d[2]);
    }
    return s;
}
__device__ __forceinline__ void split3(long long id, const Split3 &sp, int n2, int n3, int &i, int &j, int &k) {
    if (sp.small) {
        const unsigned u = (unsigned)id, q = sp.f23.div(u), r = u - q * sp.f23.d, jq = sp.f3.div(r);
        i = (int)q;
        j = (int)jq;
        k = (int)(r - jq * sp.f3.d);
    } else {
        k = (int)(id % n3);
        j = (int)((id / n3) % n2);
        i = (int)(id / ((long long)n2 * n3));
    }
}

}  // namespace lam
