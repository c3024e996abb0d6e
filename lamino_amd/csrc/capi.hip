// C ABI (include/lamino.h): argument checking, exception -> status mapping,
// and the MBIR job drivers (src/mbir_cg.cpp:47-128, src/mbir_sb.cpp:46-114,
// src/mbir3.cpp:46-135) on device-resident state.
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "capi_util.h"
#include "fft.h"
#include "nufft.h"
#include "ops.h"
#include "solvers.h"
#include "toeplitz.h"

namespace lam {

Ctx::~Ctx() {
    // Buffers first: those allocated stream-ordered are freed on this stream,
    // which must still exist.
    fft_plans.clear();
    nufft_plans.clear();
    for (DevBuf &w : work) w.release();
    red.release();
    scal.release();
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
}

namespace {

unsigned vgrid(long long n) { return grid_for((size_t)n, 256, 1u << 16); }

// y_pad[p][r + d1][c + d2] = proj[p][r][c] / mx, zero elsewhere (include/padding.h:93-131)
__global__ __launch_bounds__(256) void k_pad2d_norm(const float *__restrict__ in, float *__restrict__ out, long long P,
                                                    int R, int C, int R2, int C2, int d1, int d2, float mx) {
    const long long total = P * R2 * C2;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(id % C2), r = (int)((id / C2) % R2);
        const long long p = id / ((long long)C2 * R2);
        const int rr = r - d1, cc = c - d2;
        float v = 0.f;
        if (rr >= 0 && rr < R && cc >= 0 && cc < C) v = in[(p * R + rr) * C + cc] / mx;
        out[id] = v;
    }
}

// out[i][j][k] = in[i + o1][j + o2][k + o3]  (crop3d SYMMETRIC, include/padding.h:206-232)
__global__ __launch_bounds__(256) void k_crop3(const float *__restrict__ in, float *__restrict__ out, int m1, int m2,
                                               int m3, int n2, int n3, int o1, int o2, int o3) {
    const long long total = (long long)m1 * m2 * m3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(id % m3), j = (int)((id / m3) % m2), i = (int)(id / ((long long)m3 * m2));
        out[id] = in[(((long long)(i + o1)) * n2 + (j + o2)) * n3 + (k + o3)];
    }
}

constexpr long long kAutoLipschitzIters = 20;
constexpr uint64_t kAutoLipschitzSeed = 0x4c616d696e6f4c31ull;

// n + 2*floor(n*f/2) with f in float (include/padding.h:98, src/mbir_sb.cpp:51-55)
long long pad_n(long long n, float f) { return n + 2 * ((long long)((float)n * f) / 2); }

using clk = std::chrono::steady_clock;

}  // namespace

// Runs one reconstruction job end to end on the context's stream.
void mbir_run(Ctx &cx, const lam_dataset *ds, int nds, const long long rd[3], const lam_recon_params &pr, float *out3,
              lam_recon_stats *stats, MbirMode mode, bool transpose120) {
    LAM_REQUIRE(nds >= 1, LAM_ERR_ARG, "MBIR needs at least one dataset");
    LAM_REQUIRE(pr.method >= 1 && pr.method <= 3, LAM_ERR_ARG, "unknown reconstruction method");
    const auto t0 = clk::now();
    const float PAD = (float)pr.pad_factor;
    // MBIR1 pads the detector by (PAD-1) and MBIR2/3 by PAD (src/mbir_cg.cpp:53,83 vs src/mbir_sb.cpp:51,78)
    const float dpad = pr.method == 1 ? (float)((double)PAD - 1.0) : PAD;
    const float vpad = pr.method == 1 ? dpad : (float)(PAD - 1.0f);
    long long od[3];
    for (int a = 0; a < 3; ++a) od[a] = pad_n(rd[a], vpad);
    const long long n = od[0] * od[1] * od[2];
    const double eps = pr.eps > 0 ? pr.eps : 1e-5;

    Vol3Buf yT;
    yT.alloc(n);
    LAM_HIP(hipMemsetAsync(yT.buf.ptr, 0, yT.buf.bytes, cx.stream));
    std::vector<std::unique_ptr<Grid>> grids;
    DevBuf raw, padded;
    for (int g = 0; g < nds; ++g) {
        const lam_dataset &D = ds[g];
        need(D.proj, "dataset.proj");
        need(D.alpha, "dataset.alpha");
        LAM_REQUIRE(D.nproj >= 1 && D.nrows >= 1 && D.ncols >= 1, LAM_ERR_ARG, "invalid dataset shape");
        const long long np = D.nproj * D.nrows * D.ncols;
        raw.ensure(sizeof(float) * np);
        LAM_HIP(hipMemcpyAsync(raw.ptr, D.proj, sizeof(float) * np, hipMemcpyHostToDevice, cx.stream));
        const float mx = max_value(cx, np, raw.as<float>());
        LAM_REQUIRE(mx != 0.f, LAM_ERR_DIVZERO, "Division by zero in Array<T>::operator/=");
        const long long R2 = pad_n(D.nrows, dpad), C2 = pad_n(D.ncols, dpad);
        padded.ensure(sizeof(float) * D.nproj * R2 * C2);
        hipLaunchKernelGGL(k_pad2d_norm, dim3(vgrid(D.nproj * R2 * C2)), dim3(256), 0, cx.stream, raw.as<float>(),
                           padded.as<float>(), D.nproj, (int)D.nrows, (int)D.ncols, (int)R2, (int)C2,
                           (int)((R2 - D.nrows) / 2), (int)((C2 - D.ncols) / 2), mx);
        check_launch();
        std::vector<double> gam(D.nproj);
        for (long long p = 0; p < D.nproj; ++p) gam[p] = D.gamma ? D.gamma[p] : D.gamma_const;
        grids.push_back(make_grid(cx, D.alpha, gam.data(), D.nproj, R2, C2));
        // the graph form picks its spreading tiles by measurement, once per shape, on an
        // evenly spaced subset of the projections (same planes' density, fewer of them)
        if (mode == MbirMode::Graph && !bins_tuned(cx, *grids.back(), od, eps)) {
            constexpr long long kTuneProjections = 64;
            const long long Pt = std::min<long long>(D.nproj, kTuneProjections);
            std::vector<double> ta(Pt), tg(Pt);
            for (long long k = 0; k < Pt; ++k) {
                const long long i = k * D.nproj / Pt;
                ta[k] = D.alpha[i];
                tg[k] = gam[i];
            }
            auto sub = make_grid(cx, ta.data(), tg.data(), Pt, R2, C2);
            autotune_bins(cx, *sub, od, eps, nullptr);
        }
        float *ys[3] = {yT.v().c[0], yT.v().c[1], yT.v().c[2]};
        adjoint(cx, *grids.back(), od, eps, padded.as<float>(), ys, true);
    }
    raw.release();
    padded.release();
    std::vector<Grid *> gp;
    for (auto &g : grids) gp.push_back(g.get());
    auto toep = toep_create(cx, gp.data(), nds, od, pr.method == 3);
    LAM_HIP(hipStreamSynchronize(cx.stream));
    const auto t1 = clk::now();

    Vol3Buf x;
    x.alloc(n);
    LAM_HIP(hipMemsetAsync(x.buf.ptr, 0, x.buf.bytes, cx.stream));
    CgStats st;
    long long outer = 0;
    long long normal_ops = -1;
    double lip = 0;
    if (pr.method == 1) {
        OpFn A = [&](CVol3 u, Vol3 y) { toep_apply(*toep, u, y, 0.f); };
        if (mode == MbirMode::Graph || mode == MbirMode::Async) {
            cg_graph_solve(cx, od, A, yT.cv(), x.v(), pr.max_iters, pr.tol, mode == MbirMode::Graph, &st);
            normal_ops = st.applies;
        } else {
            cg_solve(cx, od, A, yT.cv(), x.cv(), x.v(), pr.max_iters, pr.tol, &st);
        }
        outer = 1;
    } else if (pr.method == 2) {
        const float mu = (float)pr.mu;
        OpFn A = [&](CVol3 u, Vol3 y) { toep_apply(*toep, u, y, mu); };
        if (mode == MbirMode::Graph || mode == MbirMode::Async) {
            split_bregman_graph(cx, od, A, yT.cv(), x.cv(), pr.lambda, pr.mu, pr.max_iters, pr.inner_iters, pr.tol,
                                pr.xtol, x.v(), mode == MbirMode::Graph, &st, &outer);
            normal_ops = st.applies;
        } else {
            split_bregman(cx, od, A, true, yT.cv(), x.cv(), pr.lambda, pr.mu, pr.max_iters, pr.inner_iters, pr.tol,
                          pr.xtol, x.v(), &st, &outer);
        }
    } else {
        const float G = (float)nds;
        const float sigma = (float)pr.sigma, p = (float)pr.p;
        OpFn grad = [&](CVol3 u, Vol3 g) {
            toep_apply(*toep, u, g, 0.f);
            // sum_g (sysmat_g(u) - yT) = N u - G yT   (src/mbir3.cpp:91-97)
            axpby3(cx, n, -G, yT.cv(), 1.f, g);
            qggmrf3(cx, od, g, u, 3, sigma, p);
        };
        double yTy = 0;
        {
            RedTerm t[3] = {{RED_SQ, yT.v().c[0], nullptr}, {RED_SQ, yT.v().c[1], nullptr}, {RED_SQ, yT.v().c[2], nullptr}};
            double s[3];
            reduce_terms(cx, n, t, 3, s);
            yTy = s[0] + s[1] + s[2];
        }
        LossFn loss = [&](CVol3 u) {
            std::vector<double> q(3 * nds);
            toep_quad(*toep, u, q.data());
            RedTerm t[3] = {{RED_DOT, u.c[0], yT.v().c[0]}, {RED_DOT, u.c[1], yT.v().c[1]}, {RED_DOT, u.c[2], yT.v().c[2]}};
            double s[3];
            reduce_terms(cx, n, t, 3, s);
            double e = 0;
            for (int g = 0; g < nds; ++g)
                for (int c = 0; c < 3; ++c) e += std::sqrt(q[3 * g + c] - 2.0 * s[c] + yTy);
            return e;
        };
        // src/mbir3.cpp:123-125 passes mu as the Lipschitz constant, so the step is 0.1/mu
        // whatever the operator's norm; that diverges once the normal operator's largest
        // eigenvalue exceeds 20 mu (3 x 360 projections put it in the thousands).  mu <= 0
        // asks for the normal operator's constant instead: an FP64 power iteration
        // (src/lipschitz.cpp:30-53, 20 iterations, a lower bound that converges from below).
        double L = pr.mu;
        if (L <= 0) {
            OpFn N = [&](CVol3 u, Vol3 y) { toep_apply(*toep, u, y, 0.f); };
            L = lipschitz(cx, od, N, kAutoLipschitzIters, 0.0, kAutoLipschitzSeed);
            LAM_REQUIRE(L > 0, LAM_ERR_ARG, "MBIR3: the normal operator vanishes (Lipschitz estimate 0)");
        }
        lip = L;
        long long calls = 0;
        if (mode == MbirMode::Graph || mode == MbirMode::Async) {
            const long long ops0 = toep->applies;
            LossTermsFn terms = [&](CVol3 u, double *q, double *sd, double *qpart, double *rpart) {
                toep_quad_dev(*toep, u, q, qpart);
                RedTerm t[3] = {{RED_DOT, u.c[0], yT.v().c[0]}, {RED_DOT, u.c[1], yT.v().c[1]},
                                {RED_DOT, u.c[2], yT.v().c[2]}};
                reduce_terms_dev(cx, n, t, 3, sd, rpart);
            };
            nagopt_graph(cx, od, grad, terms, nds, yTy, x.v(), pr.max_iters, L, pr.tol, pr.xtol, pr.inner_iters,
                         mode == MbirMode::Graph, &outer, &calls);
            // the replays do not pass through the host's operator count
            normal_ops = ops0 + calls;
        } else {
            nagopt(cx, od, grad, loss, x.v(), pr.max_iters, L, pr.tol, pr.xtol, pr.inner_iters, &outer, &calls);
        }
    }
    // crop to the requested dims and download
    const long long m = rd[0] * rd[1] * rd[2];
    DevBuf crop(sizeof(float) * m), tr;
    if (transpose120) tr.alloc(sizeof(float) * m);
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_crop3, dim3(vgrid(m)), dim3(256), 0, cx.stream, x.v().c[c], crop.as<float>(), (int)rd[0],
                           (int)rd[1], (int)rd[2], (int)od[1], (int)od[2], (int)((od[0] - rd[0]) / 2),
                           (int)((od[1] - rd[1]) / 2), (int)((od[2] - rd[2]) / 2));
        check_launch();
        const float *res = crop.as<float>();
        if (transpose120) {
            transpose120_dev(cx, rd, crop.as<float>(), tr.as<float>());
            res = tr.as<float>();
        }
        LAM_HIP(hipMemcpyAsync(out3 + c * m, res, sizeof(float) * m, hipMemcpyDeviceToHost, cx.stream));
    }
    LAM_HIP(hipStreamSynchronize(cx.stream));
    const auto t2 = clk::now();
    if (stats) {
        stats->setup_s = std::chrono::duration<double>(t1 - t0).count();
        stats->solve_s = std::chrono::duration<double>(t2 - t1).count();
        stats->normal_ops = normal_ops >= 0 ? normal_ops : toep->applies;
        stats->cg_iters = st.iters;
        stats->outer_iters = outer;
        stats->lipschitz = lip;
    }
}

// Upper bound of the device bytes mbir_run allocates for a job: every buffer
// of the setup (projections, point bins, fine grid, detector spectra) counted
// as if alive together with the Toeplitz operator and the solver state.
size_t mbir_footprint(const lam_dataset *ds, int nds, const long long rd[3], const lam_recon_params &pr) {
    LAM_REQUIRE(nds >= 1, LAM_ERR_ARG, "MBIR needs at least one dataset");
    LAM_REQUIRE(pr.method >= 1 && pr.method <= 3, LAM_ERR_ARG, "unknown reconstruction method");
    const float PAD = (float)pr.pad_factor;
    const float dpad = pr.method == 1 ? (float)((double)PAD - 1.0) : PAD;
    const float vpad = pr.method == 1 ? dpad : (float)(PAD - 1.0f);
    long long od[3];
    for (int a = 0; a < 3; ++a) od[a] = pad_n(rd[a], vpad);
    const size_t n = (size_t)(od[0] * od[1] * od[2]);
    const double eps = pr.eps > 0 ? pr.eps : 1e-5;
    size_t nf = 1, lt = 1, tiles = 1;
    for (int a = 0; a < 3; ++a) {
        const size_t f = nufft_fine_len(od[a], eps);
        nf *= f;
        tiles *= (f + 1) / 2;                       // bins are never smaller than 2 cells
        lt *= (size_t)toep_len(od[a]);
    }
    size_t raw = 0, padded = 0, spectra = 0, bintmp = 0, bins = 0, rows_slab = 0;
    for (int g = 0; g < nds; ++g) {
        const lam_dataset &D = ds[g];
        const size_t R2 = (size_t)pad_n(D.nrows, dpad), C2 = (size_t)pad_n(D.ncols, dpad);
        const size_t P = (size_t)D.nproj, M = P * R2 * C2, half = P * ((R2 * C2 + 1) / 2);
        raw = std::max(raw, 4 * P * (size_t)D.nrows * (size_t)D.ncols);
        padded = std::max(padded, 4 * M);
        spectra = std::max(spectra, 8 * M);
        // sort keys, the bucket-scattered (cell, index) pairs, bucket offsets and cursors
        bintmp = std::max(bintmp, 16 * half + 12 * tiles + ((size_t)16 << 20));
        // sorted ids, chunk list, geometry; twice: the autotuner holds two tile shapes at once
        bins += 2 * (4 * half + 16 * (half / 1024 + tiles)) + 128 * P;
        // the row-factorised projectors (grids sharing the tilt; counted for every grid)
        const RowBytes rb = row_path_bytes(od, eps, (long long)P, (long long)R2, (long long)C2);
        rows_slab = std::max(rows_slab, rb.slab);
        bintmp = std::max(bintmp, rb.setup);
        bins += rb.kept;
    }
    const size_t fine = 24 * nf;                                   // 3 complex components
    const size_t toep = 24 * lt * (1 + (pr.method == 3 ? (size_t)nds : 0)) + 32 * lt;
    size_t solve = 12 * n;                                         // x
    if (pr.method == 1) solve += 48 * n + 8 * n;                   // r z p Ap + preconditioner
    if (pr.method == 2) solve += 24 * n + 72 * n + 48 * n + 8 * n; // xold rhs, d b, PCG
    if (pr.method == 3) solve += 36 * n;                           // xold y g
    const size_t crop = 8 * (size_t)(rd[0] * rd[1] * rd[2]);      // cropped + transposed component
    const size_t slack = 64ull << 20;                              // plans, reductions, twiddles
    return 12 * n + raw + padded + spectra + bintmp + bins + rows_slab + fine + toep + solve + crop + slack;
}

}  // namespace lam

using namespace lam;

namespace {
struct CbOp {
    lam_op_fn fn;
    void *user;
};
}  // namespace

extern "C" {

const char *lam_last_error(void) { return g_last_error.c_str(); }

lam_status lam_ctx_create(int device, lam_ctx **out) {
    return guard([&] {
        need(out, "out");
        int count = 0;
        LAM_HIP(hipGetDeviceCount(&count));
        LAM_REQUIRE(device >= 0 && device < count, LAM_ERR_ARG, "invalid device ordinal");
        LAM_HIP(hipSetDevice(device));
        auto c = std::make_unique<lam_ctx>();
        c->impl.device = device;
        LAM_HIP(hipStreamCreateWithFlags(&c->impl.stream, hipStreamNonBlocking));
        c->impl.own_stream = true;
        *out = c.release();
    });
}

lam_status lam_ctx_destroy(lam_ctx *ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->impl.device);
        (void)hipStreamSynchronize(ctx->impl.stream);
        delete ctx;
    });
}

lam_status lam_ctx_set_stream(lam_ctx *ctx, void *s) {
    return guard([&] {
        need(ctx, "ctx");
        Ctx &c = ctx->impl;
        if (s) {
            if (c.own_stream && c.stream) LAM_HIP(hipStreamDestroy(c.stream));
            c.stream = static_cast<hipStream_t>(s);
            c.own_stream = false;
        } else if (!c.own_stream) {
            LAM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            c.own_stream = true;
        }
    });
}

void *lam_ctx_stream(lam_ctx *ctx) { return ctx ? (void *)ctx->impl.stream : nullptr; }

lam_status lam_ctx_sync(lam_ctx *ctx) {
    return guard([&] {
        need(ctx, "ctx");
        LAM_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

lam_status lam_malloc(lam_ctx *ctx, size_t bytes, void **ptr) {
    return guard([&] {
        need(ctx, "ctx");
        need(ptr, "ptr");
        *ptr = nullptr;
        if (bytes) {
            hipError_t e = hipMalloc(ptr, bytes);
            if (e == hipErrorOutOfMemory) throw Error(LAM_ERR_NOMEM, "device allocation failed");
            LAM_HIP(e);
        }
    });
}

lam_status lam_free(lam_ctx *ctx, void *ptr) {
    return guard([&] {
        need(ctx, "ctx");
        if (ptr) LAM_HIP(hipFree(ptr));
    });
}

lam_status lam_upload(lam_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard([&] {
        need(ctx, "ctx");
        if (!bytes) return;
        need(dst, "dst"); need(src, "src");
        LAM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->impl.stream));
        LAM_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

lam_status lam_download(lam_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard([&] {
        need(ctx, "ctx");
        if (!bytes) return;
        need(dst, "dst"); need(src, "src");
        LAM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->impl.stream));
        LAM_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

lam_status lam_memset(lam_ctx *ctx, void *dst, int v, size_t bytes) {
    return guard([&] {
        need(ctx, "ctx");
        if (!bytes) return;
        need(dst, "dst");
        LAM_HIP(hipMemsetAsync(dst, v, bytes, ctx->impl.stream));
    });
}

lam_status lam_copy_device(lam_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guard([&] {
        need(ctx, "ctx");
        if (!bytes) return;
        need(dst, "dst"); need(src, "src");
        LAM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->impl.stream));
    });
}

lam_status lam_grid_polar(lam_ctx *ctx, const double *theta, int64_t nproj, int64_t nrows, int64_t ncols, double gamma,
                          lam_grid **out) {
    return guard([&] {
        need(ctx, "ctx"); need(theta, "theta"); need(out, "out");
        std::vector<double> gam((size_t)std::max<int64_t>(nproj, 0), gamma);
        auto g = std::make_unique<lam_grid>();
        g->impl = make_grid(ctx->impl, theta, gam.data(), nproj, nrows, ncols);
        g->impl->literal = true;
        g->impl->gamma_const = gamma;
        *out = g.release();
    });
}

lam_status lam_grid_polar_pp(lam_ctx *ctx, const double *alpha, const double *gamma, int64_t nproj, int64_t nrows,
                             int64_t ncols, lam_grid **out) {
    return guard([&] {
        need(ctx, "ctx"); need(alpha, "alpha"); need(gamma, "gamma"); need(out, "out");
        auto g = std::make_unique<lam_grid>();
        g->impl = make_grid(ctx->impl, alpha, gamma, nproj, nrows, ncols);
        *out = g.release();
    });
}

lam_status lam_grid_destroy(lam_grid *grid) {
    return guard([&] { delete grid; });
}

lam_status lam_grid_shape(const lam_grid *grid, int64_t *P, int64_t *R, int64_t *C) {
    return guard([&] {
        need(grid, "grid");
        if (P) *P = grid->impl->P;
        if (R) *R = grid->impl->R;
        if (C) *C = grid->impl->C;
    });
}

lam_status lam_grid_coords_host(const lam_grid *grid, float *x, float *y, float *z) {
    return guard([&] {
        need(grid, "grid"); need(x, "x"); need(y, "y"); need(z, "z");
        const Grid &g = *grid->impl;
        const float dX = (float)((2 * M_PI) / (float)g.C);
        const float dY = (float)((2 * M_PI) / (float)g.R);
        for (long long i = 0; i < g.P; ++i) {
            const float th = (float)g.alpha[i], ga = (float)g.gamma[i];
            for (long long j = 0; j < g.R; ++j)
                for (long long k = 0; k < g.C; ++k) {
                    const float qX = (float)(k * dX - M_PI + dX / 2);
                    const float qY = (float)(j * dY - M_PI + dY / 2);
                    const size_t id = (size_t)((i * g.R + j) * g.C + k);
                    z[id] = -qX * std::sin(ga) + qY * std::cos(th) * std::cos(ga);
                    y[id] = -qY * std::sin(th);
                    x[id] = qX * std::cos(ga) + qY * std::cos(th) * std::sin(ga);
                }
        }
    });
}

lam_status lam_nufft3d1(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const void *c, void *f) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid"); need(c, "c"); need(f, "f");
        long long d[3];
        dims_of(dims, d);
        nufft1_complex(ctx->impl, *grid->impl, d, eps, static_cast<const float2 *>(c), static_cast<float2 *>(f));
    });
}

lam_status lam_nufft3d2(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, void *c, const void *f) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid"); need(c, "c"); need(f, "f");
        long long d[3];
        dims_of(dims, d);
        nufft2_complex(ctx->impl, *grid->impl, d, eps, static_cast<float2 *>(c), static_cast<const float2 *>(f));
    });
}

lam_status lam_forward(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *m0,
                       const float *m1, const float *m2, float *proj) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2"); need(proj, "proj");
        long long d[3];
        dims_of(dims, d);
        const float *m[3] = {m0, m1, m2};
        forward(ctx->impl, *grid->impl, d, eps, m, proj);
    });
}

static lam_status adjoint_impl(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *proj,
                               float *m0, float *m1, float *m2, bool acc) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid"); need(proj, "proj"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2");
        long long d[3];
        dims_of(dims, d);
        float *m[3] = {m0, m1, m2};
        adjoint(ctx->impl, *grid->impl, d, eps, proj, m, acc);
    });
}

lam_status lam_adjoint(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *proj, float *m0,
                       float *m1, float *m2) {
    return adjoint_impl(ctx, grid, dims, eps, proj, m0, m1, m2, false);
}

lam_status lam_adjoint_accumulate(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *proj,
                                  float *m0, float *m1, float *m2) {
    return adjoint_impl(ctx, grid, dims, eps, proj, m0, m1, m2, true);
}

static std::vector<Grid *> grid_list(lam_grid *const *grids, int ngrids) {
    LAM_REQUIRE(grids && ngrids >= 1, LAM_ERR_ARG, "no grids");
    std::vector<Grid *> gp(ngrids);
    for (int k = 0; k < ngrids; ++k) {
        LAM_REQUIRE(grids[k] && grids[k]->impl, LAM_ERR_ARG, "null grid");
        gp[k] = grids[k]->impl.get();
    }
    return gp;
}

lam_status lam_forward_multi(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3], double eps,
                             const float *m0, const float *m1, const float *m2, float *const *proj) {
    return guard([&] {
        need(ctx, "ctx"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2"); need(proj, "proj");
        const auto gp = grid_list(grids, ngrids);
        for (int k = 0; k < ngrids; ++k) need(proj[k], "proj[k]");
        long long d[3];
        dims_of(dims, d);
        const float *m[3] = {m0, m1, m2};
        forward_multi(ctx->impl, gp.data(), ngrids, d, eps, m, proj);
    });
}

lam_status lam_adjoint_multi(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3], double eps,
                             const float *const *proj, float *m0, float *m1, float *m2, int accumulate) {
    return guard([&] {
        need(ctx, "ctx"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2"); need(proj, "proj");
        const auto gp = grid_list(grids, ngrids);
        for (int k = 0; k < ngrids; ++k) need(proj[k], "proj[k]");
        long long d[3];
        dims_of(dims, d);
        float *m[3] = {m0, m1, m2};
        adjoint_multi(ctx->impl, gp.data(), ngrids, d, eps, proj, m, accumulate != 0);
    });
}

lam_status lam_sysmat(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *const x[3],
                      float *const y[3]) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid"); need(x, "x"); need(y, "y");
        long long d[3];
        dims_of(dims, d);
        sysmat_nufft(ctx->impl, *grid->impl, d, eps, x, y);
    });
}

lam_status lam_toep_create(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3], int keep,
                           lam_toep **out) {
    return guard([&] {
        need(ctx, "ctx"); need(grids, "grids"); need(out, "out");
        long long d[3];
        dims_of(dims, d);
        std::vector<Grid *> gp;
        for (int g = 0; g < ngrids; ++g) {
            need(grids[g], "grid");
            gp.push_back(grids[g]->impl.get());
        }
        auto t = std::make_unique<lam_toep>();
        t->impl = toep_create(ctx->impl, gp.data(), ngrids, d, keep != 0);
        *out = t.release();
    });
}

lam_status lam_toep_destroy(lam_toep *op) {
    return guard([&] { delete op; });
}

lam_status lam_toep_apply(lam_toep *op, const float *const x[3], float *const y[3], float mu) {
    return guard([&] {
        need(op, "op"); need(x, "x"); need(y, "y");
        toep_apply(*op->impl, CVol3{{x[0], x[1], x[2]}}, Vol3{{y[0], y[1], y[2]}}, mu);
    });
}

lam_status lam_toep_quad(lam_toep *op, const float *const x[3], double *q) {
    return guard([&] {
        need(op, "op"); need(x, "x"); need(q, "q");
        toep_quad(*op->impl, CVol3{{x[0], x[1], x[2]}}, q);
    });
}

lam_status lam_fft2(lam_ctx *ctx, const int64_t dims[3], int sign, int shifted, const void *in, void *out) {
    return guard([&] {
        need(ctx, "ctx"); need(in, "in"); need(out, "out");
        LAM_REQUIRE(sign == -1 || sign == 1, LAM_ERR_ARG, "sign must be -1 or +1");
        long long d[3];
        dims_of(dims, d);
        Ctx &c = ctx->impl;
        // pass 1 along n3 (in -> out), pass 2 along n2 (in place on out)
        AxisPass a;
        a.in = in; a.is1 = d[1] * d[2]; a.is2 = d[2]; a.isn = 1;
        a.out = out; a.os1 = d[1] * d[2]; a.os2 = d[2]; a.osn = 1;
        a.c1 = (int)d[0]; a.c2 = (int)d[1]; a.ci = 1;
        if (shifted) { a.rot_in = (int)(d[2] / 2); a.rot_out = (int)(d[2] / 2); }
        run_axis(c, get_fft_plan(c, (int)d[2], sign), a);
        AxisPass b;
        b.in = out; b.is1 = d[1] * d[2]; b.isi = 1; b.isn = d[2];
        b.out = out; b.os1 = d[1] * d[2]; b.osi = 1; b.osn = d[2];
        b.c1 = (int)d[0]; b.c2 = 1; b.ci = (int)d[2];
        if (shifted) { b.rot_in = (int)(d[1] / 2); b.rot_out = (int)(d[1] / 2); }
        run_axis(c, get_fft_plan(c, (int)d[1], sign), b);
    });
}

lam_status lam_precond(lam_ctx *ctx, const int64_t dims[3], const float *in, float *out) {
    return guard([&] {
        need(ctx, "ctx"); need(in, "in"); need(out, "out");
        long long d[3];
        dims_of(dims, d);
        precond_apply(ctx->impl, d, in, nullptr, out, nullptr);
    });
}

lam_status lam_grad_u(lam_ctx *ctx, const int64_t dims[3], const float *u, float *g0, float *g1, float *g2) {
    return guard([&] {
        need(ctx, "ctx"); need(u, "u"); need(g0, "g0"); need(g1, "g1"); need(g2, "g2");
        long long d[3];
        dims_of(dims, d);
        grad_u(ctx->impl, d, u, g0, g1, g2);
    });
}

lam_status lam_divergence(lam_ctx *ctx, const int64_t dims[3], const float *v0, const float *v1, const float *v2,
                          float *out) {
    return guard([&] {
        need(ctx, "ctx"); need(v0, "v0"); need(v1, "v1"); need(v2, "v2"); need(out, "out");
        long long d[3];
        dims_of(dims, d);
        divergence(ctx->impl, d, v0, v1, v2, out);
    });
}

lam_status lam_laplacian(lam_ctx *ctx, const int64_t dims[3], const float *u, float *out) {
    return guard([&] {
        need(ctx, "ctx"); need(u, "u"); need(out, "out");
        long long d[3];
        dims_of(dims, d);
        laplacian(ctx->impl, d, u, out);
    });
}

lam_status lam_qggmrf(lam_ctx *ctx, const int64_t dims[3], float *g, const float *x, float sigma, float p) {
    return guard([&] {
        need(ctx, "ctx"); need(g, "g"); need(x, "x");
        long long d[3];
        dims_of(dims, d);
        qggmrf(ctx->impl, d, g, x, sigma, p);
    });
}

lam_status lam_dot(lam_ctx *ctx, int64_t n, const float *a, const float *b, double *out) {
    return guard([&] {
        need(ctx, "ctx"); need(a, "a"); need(b, "b"); need(out, "out");
        RedTerm t{RED_DOT, a, b};
        reduce_terms(ctx->impl, n, &t, 1, out);
    });
}

lam_status lam_dot_noalloc(lam_ctx *ctx, int64_t n, const float *a, const float *b, double *out) {
    return guard([&] {
        need(ctx, "ctx"); need(a, "a"); need(b, "b"); need(out, "out");
        *out = dot_noalloc(ctx->impl, n, a, b);
    });
}

// src/forward.cpp:123-183: transpose {2, 0, 1}, pad3d by sqrt(2) - 1 (float factor,
// include/padding.h:168-204), forward projection on PolarGrid(angles, padded(rows),
// padded(cols), gamma), crop2d back to rows x cols.
lam_status lam_forward_simulate(lam_ctx *ctx, const float *m0, const float *m1, const float *m2,
                                const int64_t dims[3], const double *angles, int64_t nangles, double gamma,
                                int64_t rows, int64_t cols, double eps, float *proj_host) {
    return guard([&] {
        need(ctx, "ctx"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2"); need(angles, "angles");
        need(proj_host, "proj_host");
        long long d[3];
        dims_of(dims, d);
        LAM_REQUIRE(nangles >= 1 && rows >= 1 && cols >= 1, LAM_ERR_ARG, "invalid projection shape");
        Ctx &cx = ctx->impl;
        constexpr double kPadding = 1.41421356237;
        const float f = (float)(kPadding - 1);
        const long long td[3] = {d[2], d[0], d[1]};   // transposed {2, 0, 1}
        long long pd[3];
        for (int a = 0; a < 3; ++a) pd[a] = td[a] + 2 * ((long long)((float)td[a] * f) / 2);
        auto padded = [&](long long n) { return n + 2 * (long long)((double)n * (kPadding - 1) / 2); };
        const long long R2 = padded(rows), C2 = padded(cols);
        const long long n = d[0] * d[1] * d[2], np = pd[0] * pd[1] * pd[2];
        DevBuf raw(sizeof(float) * n), vol(sizeof(float) * 3 * np);
        LAM_HIP(hipMemsetAsync(vol.ptr, 0, vol.bytes, cx.stream));
        const float *mh[3] = {m0, m1, m2};
        float *mv[3];
        for (int c = 0; c < 3; ++c) {
            mv[c] = vol.as<float>() + c * np;
            LAM_HIP(hipMemcpyAsync(raw.ptr, mh[c], sizeof(float) * n, hipMemcpyHostToDevice, cx.stream));
            transpose201_pad(cx, d, raw.as<float>(), pd, mv[c]);
        }
        raw.release();
        std::vector<double> gam((size_t)nangles, gamma);
        auto grid = make_grid(cx, angles, gam.data(), nangles, R2, C2);
        DevBuf proj(sizeof(float) * nangles * R2 * C2), out(sizeof(float) * nangles * rows * cols);
        const float *mc[3] = {mv[0], mv[1], mv[2]};
        forward(cx, *grid, pd, eps > 0 ? eps : 1e-5, mc, proj.as<float>());
        crop2(cx, proj.as<float>(), nangles, (int)R2, (int)C2, out.as<float>(), (int)rows, (int)cols);
        LAM_HIP(hipMemcpyAsync(proj_host, out.ptr, out.bytes, hipMemcpyDeviceToHost, cx.stream));
        LAM_HIP(hipStreamSynchronize(cx.stream));
    });
}

lam_status lam_demag(lam_ctx *ctx, const int64_t dims[3], const float *m0, const float *m1, const float *m2,
                     float *h0, float *h1, float *h2) {
    return guard([&] {
        need(ctx, "ctx"); need(m0, "m0"); need(m1, "m1"); need(m2, "m2"); need(h0, "h0"); need(h1, "h1");
        need(h2, "h2");
        long long d[3];
        dims_of(dims, d);
        Ctx &cx = ctx->impl;
        DevBuf work(sizeof(float2) * 3 * (size_t)(d[0] * d[1] * d[2]));
        const float *m[3] = {m0, m1, m2};
        float *h[3] = {h0, h1, h2};
        demag(cx, d, m, h, work.as<float2>());
        LAM_HIP(hipStreamSynchronize(cx.stream));
    });
}

lam_status lam_mask_infs_nans(lam_ctx *ctx, int64_t n, float *data) {
    return guard([&] {
        need(ctx, "ctx"); need(data, "data");
        LAM_REQUIRE(n >= 0, LAM_ERR_ARG, "n must be non-negative");
        mask_infs_nans(ctx->impl, n, data);
    });
}

lam_status lam_transpose120(lam_ctx *ctx, const int64_t dims[3], const float *in, float *out) {
    return guard([&] {
        need(ctx, "ctx"); need(in, "in"); need(out, "out");
        long long d[3];
        dims_of(dims, d);
        LAM_REQUIRE(in != out, LAM_ERR_ARG, "transpose is out of place");
        transpose120_dev(ctx->impl, d, in, out);
    });
}

lam_status lam_max(lam_ctx *ctx, int64_t n, const float *a, float *out) {
    return guard([&] {
        need(ctx, "ctx"); need(a, "a"); need(out, "out");
        *out = max_value(ctx->impl, n, a);
    });
}

lam_status lam_cgsolver(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, const float *const b[3],
                        const float *const x0[3], int64_t max_iters, double tol, float *const x[3],
                        int64_t *iters_done) {
    return guard([&] {
        need(ctx, "ctx"); need((const void *)A, "A"); need(b, "b"); need(x0, "x0"); need(x, "x");
        long long d[3];
        dims_of(dims, d);
        OpFn op = [&](CVol3 u, Vol3 y) {
            lam_status s = A(user, u.c, y.c);
            if (s != LAM_OK) throw Error(s, "operator callback failed");
        };
        CgStats st;
        cg_solve(ctx->impl, d, op, CVol3{{b[0], b[1], b[2]}}, CVol3{{x0[0], x0[1], x0[2]}}, Vol3{{x[0], x[1], x[2]}},
                 max_iters, tol, &st);
        if (iters_done) *iters_done = st.iters;
    });
}

lam_status lam_split_bregman(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, const float *const yT[3],
                             const float *const x0[3], double lambda, double mu, int64_t outer_max, int64_t inner_max,
                             double tol, double xtol, float *const x[3]) {
    return guard([&] {
        need(ctx, "ctx"); need((const void *)A, "A"); need(yT, "yT"); need(x0, "x0"); need(x, "x");
        long long d[3];
        dims_of(dims, d);
        OpFn op = [&](CVol3 u, Vol3 y) {
            lam_status s = A(user, u.c, y.c);
            if (s != LAM_OK) throw Error(s, "operator callback failed");
        };
        split_bregman(ctx->impl, d, op, false, CVol3{{yT[0], yT[1], yT[2]}}, CVol3{{x0[0], x0[1], x0[2]}}, lambda, mu,
                      outer_max, inner_max, tol, xtol, Vol3{{x[0], x[1], x[2]}}, nullptr, nullptr);
    });
}

lam_status lam_nagopt(lam_ctx *ctx, const int64_t dims[3], lam_op_fn grad, lam_loss_fn loss, void *user,
                      float *const x[3], int64_t max_iters, double L, double tol, double xtol, int64_t max_inner) {
    return guard([&] {
        need(ctx, "ctx"); need((const void *)grad, "grad"); need((const void *)loss, "loss"); need(x, "x");
        long long d[3];
        dims_of(dims, d);
        OpFn gop = [&](CVol3 u, Vol3 y) {
            lam_status s = grad(user, u.c, y.c);
            if (s != LAM_OK) throw Error(s, "gradient callback failed");
        };
        LossFn lf = [&](CVol3 u) {
            double e = 0;
            LAM_HIP(hipStreamSynchronize(ctx->impl.stream));
            lam_status s = loss(user, u.c, &e);
            if (s != LAM_OK) throw Error(s, "loss callback failed");
            return e;
        };
        nagopt(ctx->impl, d, gop, lf, Vol3{{x[0], x[1], x[2]}}, max_iters, L, tol, xtol, max_inner, nullptr, nullptr);
    });
}

lam_status lam_lipschitz(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, int64_t max_iters, double tol,
                         uint64_t seed, double *L) {
    return guard([&] {
        need(ctx, "ctx"); need((const void *)A, "A"); need(L, "L");
        long long d[3];
        dims_of(dims, d);
        OpFn op = [&](CVol3 u, Vol3 y) {
            lam_status s = A(user, u.c, y.c);
            if (s != LAM_OK) throw Error(s, "operator callback failed");
        };
        *L = lipschitz(ctx->impl, d, op, max_iters, tol, seed);
    });
}

lam_status lam_toep_op(void *user, const float *const x[3], float *const y[3]) {
    return guard([&] {
        need(user, "user"); need(x, "x"); need(y, "y");
        toep_apply(*static_cast<lam_toep *>(user)->impl, CVol3{{x[0], x[1], x[2]}}, Vol3{{y[0], y[1], y[2]}}, 0.f);
    });
}

lam_status lam_timing_enable(lam_ctx *ctx, int on) {
    return guard([&] {
        need(ctx, "ctx");
        Ctx &c = ctx->impl;
        LAM_HIP(hipStreamSynchronize(c.stream));
        c.timer.clear();
        c.timer.on = on != 0;
    });
}

lam_status lam_timing_only(lam_ctx *ctx, const char *name) {
    return guard([&] {
        need(ctx, "ctx");
        Ctx &c = ctx->impl;
        LAM_HIP(hipStreamSynchronize(c.stream));
        c.timer.only = name ? name : "";
    });
}

lam_status lam_timing_query(lam_ctx *ctx, const char *name, double *total_ms, int64_t *count) {
    return guard([&] {
        need(ctx, "ctx"); need(name, "name");
        Ctx &c = ctx->impl;
        LAM_HIP(hipStreamSynchronize(c.stream));
        double tot = 0;
        int64_t n = 0;
        auto it = c.timer.ev.find(name);
        if (it != c.timer.ev.end()) {
            for (auto &pr : it->second) {
                float ms = 0.f;
                LAM_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
                tot += ms;
                ++n;
            }
        }
        if (total_ms) *total_ms = tot;
        if (count) *count = n;
    });
}

lam_status lam_autotune_bins(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, int tile_out[3],
                             double *best_ms) {
    return guard([&] {
        need(ctx, "ctx"); need(grid, "grid");
        long long d[3];
        dims_of(dims, d);
        const auto T = autotune_bins(ctx->impl, *grid->impl, d, eps > 0 ? eps : 1e-5, best_ms);
        if (tile_out)
            for (int a = 0; a < 3; ++a) tile_out[a] = T[a];
    });
}

lam_status lam_mbir(lam_ctx *ctx, const lam_dataset *datasets, int nds, const int64_t recon_dims[3],
                    const lam_recon_params *params, float *out3, lam_recon_stats *stats) {
    return guard([&] {
        need(ctx, "ctx"); need(datasets, "datasets"); need(params, "params"); need(out3, "out3");
        long long rd[3];
        dims_of(recon_dims, rd);
        mbir_run(ctx->impl, datasets, nds, rd, *params, out3, stats, MbirMode::Plain);
    });
}

}  // extern "C"
