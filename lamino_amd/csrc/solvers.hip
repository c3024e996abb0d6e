// Device-resident PCG, split-Bregman and Nesterov solvers and the MBIR drivers.
#include <chrono>
#include <cmath>
#include <limits>
#include <random>

#include "solvers.h"

namespace lam {

namespace {

unsigned vgrid(long long n) { return grid_for((size_t)n, 256, 1u << 16); }

// x += a p ; r -= a Ap  (3 components)
__global__ __launch_bounds__(256) void k_cg_step(long long n, float alpha, Vol3 x, Vol3 r, CVol3 p, CVol3 Ap) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x.c[c][i] += p.c[c][i] * alpha;
            r.c[c][i] -= Ap.c[c][i] * alpha;
        }
    }
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_cg_dir(long long n, float beta, CVol3 z, Vol3 p) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.c[c][i] = z.c[c][i] + p.c[c][i] * beta;
    }
}

// r = b - Ax
__global__ __launch_bounds__(256) void k_sub3(long long n, CVol3 b, CVol3 Ax, Vol3 r) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r.c[c][i] = b.c[c][i] - Ax.c[c][i];
    }
}

struct Vol9 {
    float *c[9];
};

__device__ __forceinline__ bool interior(int i, int j, int k, int n1, int n2, int n3) {
    return i >= 1 && i < n1 - 1 && j >= 1 && j < n2 - 1 && k >= 1 && k < n3 - 1;
}

// rhs_i = yT_i + mu * divergence(d_i - b_i)   (src/bregman.cpp:81-84)
__global__ __launch_bounds__(256) void k_sb_rhs(CVol3 yT, Vol9 d, Vol9 b, Vol3 rhs, float mu, int n1, int n2,
                                                int n3, Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s[3] = {(long long)n2 * n3, n3, 1};
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        const bool in = interior(i, j, k, n1, n2, n3);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float dv = 0.f;
            if (in) {
                float A[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float *dd = d.c[3 * c + a], *bb = b.c[3 * c + a];
                    const float vp = dd[id + s[a]] - bb[id + s[a]];
                    const float vm = dd[id - s[a]] - bb[id - s[a]];
                    A[a] = vp - vm;
                }
                dv = (float)(-(double)A[0] / 2.0 - (double)A[1] / 2.0 - (double)A[2] / 2.0);
            }
            rhs.c[c][id] = yT.c[c][id] + dv * mu;
        }
    }
}

// isotropic shrinkage + Bregman update (src/bregman.cpp:91-120), gradients of x recomputed in place
__global__ __launch_bounds__(256) void k_sb_shrink(CVol3 x, Vol9 d, Vol9 b, float eps, float thr, int n1, int n2,
                                                   int n3, Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s[3] = {(long long)n2 * n3, n3, 1};
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        const bool in = interior(i, j, k, n1, n2, n3);
        float g[9], t[9];
        float sk = eps;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float gv = in ? (x.c[c][id + s[a]] - x.c[c][id - s[a]]) * 0.5f : 0.f;
                g[3 * c + a] = gv;
                const float tv = gv + b.c[3 * c + a][id];
                t[3 * c + a] = tv;
                sk += tv * tv;
            }
        sk = sqrtf(sk);
        const float f = fmaxf(sk - thr, 0.f);
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const float dn = f * (t[q] / sk);
            d.c[q][id] = dn;
            b.c[q][id] += g[q] - dn;
        }
    }
}

// ---- device-side PCG scalars (the graph form)
__global__ void k_cg_begin(CgScalars *s) {
    s->rs_old = s->sums[0] + s->sums[1] + s->sums[2];
    s->iters = 0;
    s->done = 0;
}

// |pAp| < 1e-10 stops before the update (src/conjgrad.cpp:73-76)
__global__ void k_cg_alpha(CgScalars *s) {
    if (s->done) return;
    const double pAp = s->sums[0] + s->sums[1] + s->sums[2];
    if (fabs(pAp) < 1e-10) {
        s->done = 1;
        return;
    }
    s->alpha = s->rs_old / pAp;
    s->iters += 1;
    s->total += 1;
}

// sqrt(rs_new) < tol stops after the update (src/conjgrad.cpp:94)
__global__ void k_cg_beta(CgScalars *s, double tol) {
    if (s->done) return;
    const double rs_new = s->sums[0] + s->sums[1] + s->sums[2];
    if (sqrt(rs_new) < tol) {
        s->done = 1;
        return;
    }
    s->beta = rs_new / s->rs_old;
    s->rs_old = rs_new;
}

__global__ __launch_bounds__(256) void k_cg_step_dev(long long n, const CgScalars *s, Vol3 x, Vol3 r, CVol3 p,
                                                     CVol3 Ap) {
    if (s->done) return;
    const float alpha = (float)s->alpha;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x.c[c][i] += p.c[c][i] * alpha;
            r.c[c][i] -= Ap.c[c][i] * alpha;
        }
    }
}

__global__ __launch_bounds__(256) void k_cg_dir_dev(long long n, const CgScalars *s, CVol3 z, Vol3 p) {
    if (s->done) return;
    const float beta = (float)s->beta;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) p.c[c][i] = z.c[c][i] + p.c[c][i] * beta;
    }
}

__global__ void k_zero_bytes(unsigned *p, long long nwords) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nwords; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// y = x + (x - xold) * beta   (src/nagopt.cpp:156-158)
__global__ __launch_bounds__(256) void k_nag_y(long long n, float beta, CVol3 x, CVol3 xold, Vol3 y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) y.c[c][i] = x.c[c][i] + (x.c[c][i] - xold.c[c][i]) * beta;
    }
}

// x = y - g * step   (src/nagopt.cpp:162)
__global__ __launch_bounds__(256) void k_nag_x(long long n, float step, CVol3 y, CVol3 g, Vol3 x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x.c[c][i] = y.c[c][i] - g.c[c][i] * step;
    }
}

}  // namespace

void cg_solve(Ctx &cx, const long long d[3], const OpFn &A, CVol3 b, CVol3 x0, Vol3 x, long long max_iters,
              double tol, CgStats *st) {
    const long long n = d[0] * d[1] * d[2];
    Vol3Buf r, z, p, Ap;
    r.alloc(n); z.alloc(n); p.alloc(n); Ap.alloc(n);
    if (x.c[0] != x0.c[0]) copy3(cx, n, x0, x);
    A(cv(x), Ap.v());
    long long applies = 1;
    hipLaunchKernelGGL(k_sub3, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, b, Ap.cv(), r.v());
    check_launch();
    auto precond3 = [&](CVol3 in, Vol3 out) {
        precond_apply(cx, d, in.c[0], in.c[1], out.c[0], out.c[1]);
        precond_apply(cx, d, in.c[2], nullptr, out.c[2], nullptr);
    };
    auto dot3 = [&](CVol3 a, CVol3 bb) {
        RedTerm t[3] = {{RED_DOT, a.c[0], bb.c[0]}, {RED_DOT, a.c[1], bb.c[1]}, {RED_DOT, a.c[2], bb.c[2]}};
        double s[3];
        reduce_terms(cx, n, t, 3, s);
        return s[0] + s[1] + s[2];
    };
    precond3(r.cv(), z.v());
    copy3(cx, n, z.cv(), p.v());
    double rs_old = dot3(z.cv(), r.cv());
    long long it = 0;
    for (; it < max_iters; ++it) {
        A(p.cv(), Ap.v());
        ++applies;
        const double pAp = dot3(p.cv(), Ap.cv());
        if (std::fabs(pAp) < 1e-10) break;
        const double alpha = rs_old / pAp;
        hipLaunchKernelGGL(k_cg_step, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (float)alpha, x, r.v(), p.cv(),
                           Ap.cv());
        check_launch();
        precond3(r.cv(), z.v());
        const double rs_new = dot3(z.cv(), r.cv());
        if (std::sqrt(rs_new) < tol) { ++it; break; }
        const double beta = rs_new / rs_old;
        hipLaunchKernelGGL(k_cg_dir, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (float)beta, z.cv(), p.v());
        check_launch();
        rs_old = rs_new;
    }
    if (st) { st->iters += it; st->applies += applies; }
}

void split_bregman(Ctx &cx, const long long d[3], const OpFn &A, bool A_has_lap, CVol3 yT, CVol3 x0, double lambda,
                   double mu, long long outer_max, long long inner_max, double tol, double xtol, Vol3 x,
                   CgStats *st, long long *outer_done) {
    const long long n = d[0] * d[1] * d[2];
    Vol3Buf xold, rhs;
    xold.alloc(n); rhs.alloc(n);
    DevBuf dd(sizeof(float) * 9 * (size_t)n), bb(sizeof(float) * 9 * (size_t)n);
    LAM_HIP(hipMemsetAsync(dd.ptr, 0, dd.bytes, cx.stream));
    LAM_HIP(hipMemsetAsync(bb.ptr, 0, bb.bytes, cx.stream));
    Vol9 D, B;
    for (int q = 0; q < 9; ++q) { D.c[q] = dd.as<float>() + q * n; B.c[q] = bb.as<float>() + q * n; }
    copy3(cx, n, x0, xold.v());
    if (x.c[0] != x0.c[0]) copy3(cx, n, x0, x);
    OpFn Ap = A;
    if (!A_has_lap) {
        Ap = [&](CVol3 u, Vol3 y) {
            A(u, y);
            add_mu_laplacian3(cx, d, u, y, (float)mu);
        };
    }
    const float eps = 1e-8f;
    const float thr = (float)lambda / (float)mu;
    long long it = 0;
    for (; it < outer_max; ++it) {
        hipLaunchKernelGGL(k_sb_rhs, dim3(vgrid(n)), dim3(256), 0, cx.stream, yT, D, B, rhs.v(), (float)mu, (int)d[0],
                           (int)d[1], (int)d[2], make_split3(d));
        check_launch();
        cg_solve(cx, d, Ap, rhs.cv(), xold.cv(), x, inner_max, tol, st);
        hipLaunchKernelGGL(k_sb_shrink, dim3(vgrid(n)), dim3(256), 0, cx.stream, cv(x), D, B, eps, thr, (int)d[0],
                           (int)d[1], (int)d[2], make_split3(d));
        check_launch();
        RedTerm t[6];
        for (int c = 0; c < 3; ++c) {
            t[c] = RedTerm{RED_DIFF2, x.c[c], xold.v().c[c]};
            t[3 + c] = RedTerm{RED_SQ, xold.v().c[c], nullptr};
        }
        double s[6];
        reduce_terms(cx, n, t, 6, s);
        double nd = 0;
        for (int c = 0; c < 3; ++c) nd += std::sqrt(s[c]) / (std::sqrt(s[3 + c]) + 1e-8);
        copy3(cx, n, cv(x), xold.v());
        if (nd < xtol) { ++it; break; }
    }
    if (outer_done) *outer_done = it;
}

void nagopt(Ctx &cx, const long long d[3], const OpFn &grad, const LossFn &loss, Vol3 x, long long max_iters,
            double lipschitz, double tol, double xtol, long long max_inner, long long *outer_done,
            long long *grad_calls) {
    const long long n = d[0] * d[1] * d[2];
    Vol3Buf xold, y, g;
    xold.alloc(n); y.alloc(n); g.alloc(n);
    copy3(cx, n, cv(x), xold.v());
    double t = 1, tnew = 1;
    double step = 0.1 / lipschitz, step_prev = step;
    long long iter = 0, calls = 0;
    for (; iter < max_iters; ++iter) {
        bool converged = false;
        // loss(x) after the line search is the last trial's ex: x has not changed since
        // (the same kernels on the same x give the same bits)
        bool have_ex = false;
        double ex_last = 0;
        for (long long inner = 0; inner < max_inner; ++inner) {
            const double beta = tnew * (1 / t - 1);
            tnew = 0.5 * (std::sqrt(std::pow(t, 4) + 4 * std::pow(t, 2)) - std::pow(t, 2));
            hipLaunchKernelGGL(k_nag_y, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (float)beta, cv(x), xold.cv(),
                               y.v());
            check_launch();
            grad(y.cv(), g.v());
            ++calls;
            hipLaunchKernelGGL(k_nag_x, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (float)step, y.cv(), g.cv(), x);
            check_launch();
            const double ex = loss(cv(x));
            have_ex = true;
            ex_last = ex;
            const double ey = loss(y.cv());
            RedTerm tg[3] = {{RED_SQ, g.v().c[0], nullptr}, {RED_SQ, g.v().c[1], nullptr}, {RED_SQ, g.v().c[2], nullptr}};
            double sg[3];
            reduce_terms(cx, n, tg, 3, sg);
            const double gy = 0.5 * step * (std::sqrt(sg[0]) + std::sqrt(sg[1]) + std::sqrt(sg[2]));
            if (ex > ey + gy) {
                step *= 0.9;
            } else {
                step = step_prev;
                t = tnew;
                RedTerm te[6];
                for (int c = 0; c < 3; ++c) {
                    te[c] = RedTerm{RED_DIFF2, x.c[c], xold.v().c[c]};
                    te[3 + c] = RedTerm{RED_SQ, x.c[c], nullptr};
                }
                double se[6];
                reduce_terms(cx, n, te, 6, se);
                const double xerr = (std::sqrt(se[0]) + std::sqrt(se[1]) + std::sqrt(se[2])) /
                                    (std::sqrt(se[3]) + std::sqrt(se[4]) + std::sqrt(se[5]));
                if (xerr < xtol) { converged = true; break; }
                copy3(cx, n, cv(x), xold.v());
                break;
            }
        }
        if (converged) { ++iter; break; }
        const double e = have_ex ? ex_last : loss(cv(x));
        if (e < tol) { ++iter; break; }
    }
    if (outer_done) *outer_done = iter;
    if (grad_calls) *grad_calls = calls;
}

double lipschitz(Ctx &cx, const long long d[3], const OpFn &A, long long max_iters, double tol, uint64_t seed) {
    const long long n = d[0] * d[1] * d[2];
    Vol3Buf x, Ax;
    x.alloc(n); Ax.alloc(n);
    std::vector<float> h(3 * (size_t)n);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (auto &v : h) v = (float)U(rng);
    LAM_HIP(hipMemcpyAsync(x.buf.ptr, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, cx.stream));
    auto norm3 = [&](CVol3 a) {
        RedTerm t[3] = {{RED_SQ, a.c[0], nullptr}, {RED_SQ, a.c[1], nullptr}, {RED_SQ, a.c[2], nullptr}};
        double s[3];
        reduce_terms(cx, n, t, 3, s);
        return std::sqrt(s[0] + s[1] + s[2]);
    };
    double nx = norm3(x.cv());
    axpby3(cx, n, (float)(1.0 / nx), x.cv(), 0.f, x.v());
    double L = 0;
    for (long long i = 0; i < max_iters; ++i) {
        A(x.cv(), Ax.v());
        const double na = norm3(Ax.cv());
        if (std::fabs(na) < 1e-12) { L = 0; break; }
        axpby3(cx, n, (float)(1.0 / na), Ax.cv(), 0.f, x.v());
        if (std::fabs(na - L) < tol) break;   // returns the previous estimate (src/lipschitz.cpp:50-51)
        L = na;
    }
    return L;
}

// ------------------------------------------------------------------ graph form
void FastCg::init(Ctx &c, const long long dims[3]) {
    cx = &c;
    for (int a = 0; a < 3; ++a) d[a] = dims[a];
    n = d[0] * d[1] * d[2];
    r.alloc(n); z.alloc(n); p.alloc(n); Ap.alloc(n);
    scal.alloc(sizeof(CgScalars));
    red.alloc(sizeof(double) * kRedPartials);
    pwork.alloc(sizeof(float2) * (size_t)n);
    hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(64), 0, c.stream, scal.as<unsigned>(),
                       (long long)(sizeof(CgScalars) / sizeof(unsigned)));
    check_launch();
}

namespace {

void fast_precond3(FastCg &w, CVol3 in, Vol3 out) {
    float2 *wk = w.pwork.as<float2>();
    precond_apply(*w.cx, w.d, in.c[0], in.c[1], out.c[0], out.c[1], wk);
    precond_apply(*w.cx, w.d, in.c[2], nullptr, out.c[2], nullptr, wk);
}

void fast_dot3(FastCg &w, CVol3 a, CVol3 b) {
    RedTerm t[3] = {{RED_DOT, a.c[0], b.c[0]}, {RED_DOT, a.c[1], b.c[1]}, {RED_DOT, a.c[2], b.c[2]}};
    reduce_terms_dev(*w.cx, w.n, t, 3, w.s()->sums, w.red.as<double>());
}

}  // namespace

void fast_cg_begin(FastCg &w, const OpFn &A, CVol3 b, CVol3 x0, Vol3 x) {
    Ctx &cx = *w.cx;
    if (x.c[0] != x0.c[0]) copy3(cx, w.n, x0, x);
    A(cv(x), w.Ap.v());
    hipLaunchKernelGGL(k_sub3, dim3(vgrid(w.n)), dim3(256), 0, cx.stream, w.n, b, w.Ap.cv(), w.r.v());
    check_launch();
    fast_precond3(w, w.r.cv(), w.z.v());
    copy3(cx, w.n, w.z.cv(), w.p.v());
    fast_dot3(w, w.z.cv(), w.r.cv());
    hipLaunchKernelGGL(k_cg_begin, dim3(1), dim3(1), 0, cx.stream, w.s());
    check_launch();
}

void fast_cg_iter(FastCg &w, const OpFn &A, Vol3 x, double tol) {
    Ctx &cx = *w.cx;
    A(w.p.cv(), w.Ap.v());
    fast_dot3(w, w.p.cv(), w.Ap.cv());
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1), 0, cx.stream, w.s());
    check_launch();
    hipLaunchKernelGGL(k_cg_step_dev, dim3(vgrid(w.n)), dim3(256), 0, cx.stream, w.n, (const CgScalars *)w.s(), x,
                       w.r.v(), w.p.cv(), w.Ap.cv());
    check_launch();
    fast_precond3(w, w.r.cv(), w.z.v());
    fast_dot3(w, w.z.cv(), w.r.cv());
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1), 0, cx.stream, w.s(), tol);
    check_launch();
    hipLaunchKernelGGL(k_cg_dir_dev, dim3(vgrid(w.n)), dim3(256), 0, cx.stream, w.n, (const CgScalars *)w.s(),
                       w.z.cv(), w.p.v());
    check_launch();
}

CgScalars fast_cg_read(FastCg &w) {
    CgScalars h;
    LAM_HIP(hipMemcpyAsync(&h, w.scal.ptr, sizeof(CgScalars), hipMemcpyDeviceToHost, w.cx->stream));
    LAM_HIP(hipStreamSynchronize(w.cx->stream));
    return h;
}

GraphStep::~GraphStep() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
}

bool GraphStep::run(Ctx &cx, bool use_graph, const std::function<void()> &body) {
    ++calls_;
    if (!use_graph || calls_ == 1) {
        body();
        return false;
    }
    if (!exec_) {
        const bool timing = cx.timer.on;
        cx.timer.on = false;   // no event records inside the graph
        {
            std::unique_lock<std::shared_mutex> lk(capture_mutex());
            LAM_HIP(hipStreamBeginCapture(cx.stream, hipStreamCaptureModeRelaxed));
            capturing_here() = true;
            try {
                body();
            } catch (...) {
                capturing_here() = false;
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cx.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                cx.timer.on = timing;
                throw;
            }
            capturing_here() = false;
            LAM_HIP(hipStreamEndCapture(cx.stream, &graph_));
        }
        cx.timer.on = timing;
        LAM_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    }
    LAM_HIP(hipGraphLaunch(exec_, cx.stream));
    return true;
}

long long cg_graph_solve(Ctx &cx, const long long d[3], const OpFn &A, CVol3 b, Vol3 x, long long max_iters,
                         double tol, bool use_graph, CgStats *st) {
    FastCg w;
    w.init(cx, d);
    fast_cg_begin(w, A, b, cv(x), x);
    GraphStep step;
    constexpr long long kCheck = 8;   // iterations between host reads of the stop flag when tol > 0
    for (long long it = 0; it < max_iters; ++it) {
        step.run(cx, use_graph, [&] { fast_cg_iter(w, A, x, tol); });
        if (tol > 0 && (it + 1) % kCheck == 0 && fast_cg_read(w).done) break;
    }
    const CgScalars s = fast_cg_read(w);
    if (st) {
        st->iters += s.total;
        st->applies += 1 + max_iters;
    }
    return s.total;
}

void split_bregman_graph(Ctx &cx, const long long d[3], const OpFn &A, CVol3 yT, CVol3 x0, double lambda, double mu,
                         long long outer_max, long long inner_max, double tol, double xtol, Vol3 x, bool use_graph,
                         CgStats *st, long long *outer_done) {
    const long long n = d[0] * d[1] * d[2];
    FastCg w;
    w.init(cx, d);
    Vol3Buf xold, rhs;
    xold.alloc(n); rhs.alloc(n);
    DevBuf dd(sizeof(float) * 9 * (size_t)n), bb(sizeof(float) * 9 * (size_t)n), xs(sizeof(double) * 8);
    for (DevBuf *z : {&dd, &bb}) fill(cx, 9 * n, z->as<float>(), 0.f);
    Vol9 D, B;
    for (int q = 0; q < 9; ++q) { D.c[q] = dd.as<float>() + q * n; B.c[q] = bb.as<float>() + q * n; }
    copy3(cx, n, x0, xold.v());
    if (x.c[0] != x0.c[0]) copy3(cx, n, x0, x);
    const float eps = 1e-8f;
    const float thr = (float)lambda / (float)mu;
    RedTerm t[6];
    for (int c = 0; c < 3; ++c) {
        t[c] = RedTerm{RED_DIFF2, x.c[c], xold.v().c[c]};
        t[3 + c] = RedTerm{RED_SQ, xold.v().c[c], nullptr};
    }
    // one outer iteration of src/bregman.cpp:75-131
    auto outer_step = [&] {
        hipLaunchKernelGGL(k_sb_rhs, dim3(vgrid(n)), dim3(256), 0, cx.stream, yT, D, B, rhs.v(), (float)mu, (int)d[0],
                           (int)d[1], (int)d[2], make_split3(d));
        check_launch();
        fast_cg_begin(w, A, rhs.cv(), xold.cv(), x);
        for (long long k = 0; k < inner_max; ++k) fast_cg_iter(w, A, x, tol);
        hipLaunchKernelGGL(k_sb_shrink, dim3(vgrid(n)), dim3(256), 0, cx.stream, cv(x), D, B, eps, thr, (int)d[0],
                           (int)d[1], (int)d[2], make_split3(d));
        check_launch();
        reduce_terms_dev(cx, n, t, 6, xs.as<double>(), w.red.as<double>());
        copy3(cx, n, cv(x), xold.v());
    };
    GraphStep step;
    long long it = 0;
    for (; it < outer_max; ++it) {
        step.run(cx, use_graph, outer_step);
        if (xtol > 0) {   // the x-change test needs the host (src/bregman.cpp:122-131)
            double s[6];
            LAM_HIP(hipMemcpyAsync(s, xs.ptr, sizeof(s), hipMemcpyDeviceToHost, cx.stream));
            LAM_HIP(hipStreamSynchronize(cx.stream));
            double nd = 0;
            for (int c = 0; c < 3; ++c) nd += std::sqrt(s[c]) / (std::sqrt(s[3 + c]) + 1e-8);
            if (nd < xtol) { ++it; break; }
        }
    }
    const CgScalars s = fast_cg_read(w);
    if (st) {
        st->iters += s.total;
        st->applies += it * (1 + inner_max);
    }
    if (outer_done) *outer_done = it;
}

// ------------------------------------------------------------------ NAG, graph form
namespace {

constexpr int kNagMaxGrids = 8;

// Device scalar block of one graph-form NAG solve.
struct NagScalars {
    double t, tnew, step, step0, beta;
    double qx[3 * kNagMaxGrids], sx[4], qy[3 * kNagMaxGrids], sy[4], sg[4], se[8];
    double ex, xerr;
    long long outer, inner, calls;
    int accepted, done;
};

// beta and tnew of the next trial (src/nagopt.cpp:151-154)
__global__ void k_nag_begin(NagScalars *s) {
    if (s->done) return;
    const double t = s->t;
    s->beta = s->tnew * (1 / t - 1);
    s->tnew = 0.5 * (sqrt(pow(t, 4.0) + 4 * pow(t, 2.0)) - pow(t, 2.0));
}

__global__ __launch_bounds__(256) void k_nag_y_dev(long long n, const NagScalars *s, CVol3 x, CVol3 xold, Vol3 y) {
    if (s->done) return;
    const float beta = (float)s->beta;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) y.c[c][i] = x.c[c][i] + (x.c[c][i] - xold.c[c][i]) * beta;
    }
}

__global__ __launch_bounds__(256) void k_nag_x_dev(long long n, const NagScalars *s, CVol3 y, CVol3 g, Vol3 x) {
    if (s->done) return;
    const float step = (float)s->step;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x.c[c][i] = y.c[c][i] - g.c[c][i] * step;
    }
}

// The Armijo test and its consequences (src/nagopt.cpp:164-201) from this trial's
// reductions: the same FP64 expressions, in the same order, as nagopt's host loop.
__global__ void k_nag_decide(NagScalars *s, int ngrids, double yTy, long long max_iters, long long max_inner,
                             double tol, double xtol) {
    if (s->done) return;
    double ex = 0, ey = 0;
    for (int g = 0; g < ngrids; ++g)
        for (int c = 0; c < 3; ++c) ex += sqrt(s->qx[3 * g + c] - 2.0 * s->sx[c] + yTy);
    for (int g = 0; g < ngrids; ++g)
        for (int c = 0; c < 3; ++c) ey += sqrt(s->qy[3 * g + c] - 2.0 * s->sy[c] + yTy);
    const double gy = 0.5 * s->step * (sqrt(s->sg[0]) + sqrt(s->sg[1]) + sqrt(s->sg[2]));
    s->calls += 1;
    s->accepted = 0;
    bool end = false;
    if (ex > ey + gy) {
        s->step *= 0.9;
        end = s->inner + 1 >= max_inner;
    } else {
        s->step = s->step0;
        s->t = s->tnew;
        const double xerr = (sqrt(s->se[0]) + sqrt(s->se[1]) + sqrt(s->se[2])) /
                            (sqrt(s->se[3]) + sqrt(s->se[4]) + sqrt(s->se[5]));
        s->xerr = xerr;
        if (xerr < xtol) {
            s->done = 1;
            s->outer += 1;
            return;
        }
        s->accepted = 1;
        end = true;
    }
    if (end) {
        // loss(x) after the line search is this trial's ex (x is unchanged since)
        s->ex = ex;
        s->inner = 0;
        s->outer += 1;
        if (ex < tol || s->outer >= max_iters) s->done = 1;
    } else {
        s->inner += 1;
    }
}

// xold = x after an accepted trial
__global__ __launch_bounds__(256) void k_nag_keep(long long n, const NagScalars *s, CVol3 x, Vol3 xold) {
    if (s->done || !s->accepted) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) xold.c[c][i] = x.c[c][i];
    }
}

}  // namespace

void nagopt_graph(Ctx &cx, const long long d[3], const OpFn &grad, const LossTermsFn &terms, int ngrids, double yTy,
                  Vol3 x, long long max_iters, double lipschitz, double tol, double xtol, long long max_inner,
                  bool use_graph, long long *outer_done, long long *grad_calls) {
    LAM_REQUIRE(ngrids >= 1 && ngrids <= kNagMaxGrids, LAM_ERR_ARG, "graph-form NAG takes 1 to 8 datasets");
    const long long n = d[0] * d[1] * d[2];
    Vol3Buf xold, y, g;
    xold.alloc(n); y.alloc(n); g.alloc(n);
    DevBuf scal(sizeof(NagScalars)), red(sizeof(double) * kRedPartials), qpart(sizeof(double) * toep_quad_partials());
    NagScalars h{};
    h.t = h.tnew = 1;
    h.step = h.step0 = 0.1 / lipschitz;
    h.done = max_iters <= 0 || max_inner <= 0;
    upload_now(cx, scal.ptr, &h, sizeof(h));
    NagScalars *s = scal.as<NagScalars>();
    copy3(cx, n, cv(x), xold.v());
    RedTerm tg[3], te[6];
    for (int c = 0; c < 3; ++c) {
        tg[c] = RedTerm{RED_SQ, g.v().c[c], nullptr};
        te[c] = RedTerm{RED_DIFF2, x.c[c], xold.v().c[c]};
        te[3 + c] = RedTerm{RED_SQ, x.c[c], nullptr};
    }
    // one trial of the line search (src/nagopt.cpp:148-201)
    auto trial = [&] {
        hipLaunchKernelGGL(k_nag_begin, dim3(1), dim3(1), 0, cx.stream, s);
        check_launch();
        hipLaunchKernelGGL(k_nag_y_dev, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (const NagScalars *)s, cv(x),
                           xold.cv(), y.v());
        check_launch();
        grad(y.cv(), g.v());
        hipLaunchKernelGGL(k_nag_x_dev, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (const NagScalars *)s, y.cv(),
                           g.cv(), x);
        check_launch();
        terms(cv(x), s->qx, s->sx, qpart.as<double>(), red.as<double>());
        terms(y.cv(), s->qy, s->sy, qpart.as<double>(), red.as<double>());
        reduce_terms_dev(cx, n, tg, 3, s->sg, red.as<double>());
        reduce_terms_dev(cx, n, te, 6, s->se, red.as<double>());
        hipLaunchKernelGGL(k_nag_decide, dim3(1), dim3(1), 0, cx.stream, s, ngrids, yTy, max_iters, max_inner, tol,
                           xtol);
        check_launch();
        hipLaunchKernelGGL(k_nag_keep, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, (const NagScalars *)s, cv(x),
                           xold.v());
        check_launch();
    };
    auto read = [&] {
        NagScalars r;
        LAM_HIP(hipMemcpyAsync(&r, scal.ptr, sizeof(r), hipMemcpyDeviceToHost, cx.stream));
        LAM_HIP(hipStreamSynchronize(cx.stream));
        return r;
    };
    GraphStep step;
    // Every trial advances the outer count by at most one, so max_iters - outer trials
    // never run past the iteration limit; a tolerance can stop earlier, so with one the
    // host reads the stop flag every kCheck trials.
    constexpr long long kCheck = 4;
    const long long max_trials = h.done ? 0 : max_iters * max_inner;
    long long trials = 0, outer_read = 0;
    while (trials < max_trials) {
        long long batch = max_iters - outer_read;
        if (tol > 0 || xtol > 0) batch = std::min(batch, kCheck);
        batch = std::min(std::max(batch, 1ll), max_trials - trials);
        for (long long b = 0; b < batch; ++b, ++trials) step.run(cx, use_graph, trial);
        const NagScalars r = read();
        if (r.done) break;
        outer_read = r.outer;
    }
    const NagScalars r = read();
    if (outer_done) *outer_done = r.outer;
    if (grad_calls) *grad_calls = r.calls;
}

}  // namespace lam
