// Polar-grid NUFFT on gfx950: deapodise + zero-pad, fine-grid FFT,
// exponential-of-semicircle interpolation / spreading with the point
// coordinates generated analytically per (projection, row, col).
//
// For real volumes the plane data are Hermitian: the point (R-1-j, C-1-k) is
// the negative of (j, k), so c(-q) = conj(c(q)).  The projector kernels
// therefore visit only the first half of every plane:
//   - interpolation writes c(q) and conj(c(q)) to the mirrored slot;
//   - spreading of any plane data followed by Re() equals spreading
//     c(q) + conj(c(-q)) over the half set (exact, no symmetry assumed).
#include <algorithm>
#include <cmath>

#include "nufft.h"

namespace lam {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ int wrapi(int i, int n) {
    if (i < 0) i += n;
    else if (i >= n) i -= n;
    return i;
}

// ES kernel weights for one axis: cells i0 .. i0+W-1 around fine coordinate t.
template <int W>
__device__ __forceinline__ void es_weights(double t, int nf, float beta, int &i0, float *wts) {
    const double fl = floor(t / (double)nf);
    t -= fl * (double)nf;
    const double s = ceil(t - 0.5 * W);
    i0 = (int)s;
    const float z0 = (float)((s - t) * (2.0 / W));
    const float dz = 2.0f / (float)W;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const float z = z0 + dz * (float)m;
        const float r = 1.0f - z * z;
        wts[m] = r > 0.f ? __expf(beta * (sqrtf(r) - 1.0f)) : 0.f;
    }
}

struct PointFrame {
    double dX, dY, k0, j0, sx, sy, sz;  // qX step, qY step, centres, fine scales nf/(2 pi)
};

__device__ __forceinline__ void point_coords(const Geo &g, const PointFrame &f, int j, int k, double &tx, double &ty,
                                             double &tz) {
    const double qX = ((double)k - f.k0) * f.dX;
    const double qY = ((double)j - f.j0) * f.dY;
    tx = (qX * g.ux + qY * g.vx) * f.sx;
    ty = (qY * g.vy) * f.sy;
    tz = (qX * g.uz + qY * g.vz) * f.sz;
}

// ------------------------------------------------------------------ deapodise + pad
// fine[g][c] = vol_c[i] * corr, where array index i <-> mode i - n/2 <-> fine index mode mod nf.
template <int NC, bool REAL>
__global__ __launch_bounds__(256) void k_deapod_pad(float2 *__restrict__ fine, const float *__restrict__ r0,
                                                    const float *__restrict__ r1, const float *__restrict__ r2,
                                                    const float2 *__restrict__ cin, int n1, int n2, int n3,
                                                    int nfz, int nfy, int nfx, const float *__restrict__ cz,
                                                    const float *__restrict__ cy, const float *__restrict__ cxr) {
    const long long total = (long long)nfz * nfy * nfx;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int gx = (int)(id % nfx);
        const int gy = (int)((id / nfx) % nfy);
        const int gz = (int)(id / ((long long)nfx * nfy));
        const int hz = n1 / 2, hy = n2 / 2, hx = n3 / 2;
        int iz = gz < n1 - hz ? gz + hz : (gz >= nfz - hz ? gz - nfz + hz : -1);
        int iy = gy < n2 - hy ? gy + hy : (gy >= nfy - hy ? gy - nfy + hy : -1);
        int ix = gx < n3 - hx ? gx + hx : (gx >= nfx - hx ? gx - nfx + hx : -1);
        float2 v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = make_float2(0.f, 0.f);
        if (iz >= 0 && iy >= 0 && ix >= 0) {
            const float cf = cz[iz] * cy[iy] * cxr[ix];
            const long long vi = ((long long)iz * n2 + iy) * n3 + ix;
            if (REAL) {
                v[0].x = r0[vi] * cf;
                if (NC == 3) {
                    v[1].x = r1[vi] * cf;
                    v[2].x = r2[vi] * cf;
                }
            } else {
                v[0] = cscale(cin[vi], cf);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) fine[id * NC + c] = v[c];
    }
}

// ------------------------------------------------------------------ crop + deapodise
// REAL: out_c[i] (+)= Re(fine[g][c] * corr) * scale; else complex out.
template <int NC, bool REAL>
__global__ __launch_bounds__(256) void k_crop_deapod(const float2 *__restrict__ fine, float *__restrict__ o0,
                                                     float *__restrict__ o1, float *__restrict__ o2,
                                                     float2 *__restrict__ cout, int n1, int n2, int n3, int nfz,
                                                     int nfy, int nfx, const float *__restrict__ cz,
                                                     const float *__restrict__ cy, const float *__restrict__ cxr,
                                                     float scale, int accumulate) {
    const long long total = (long long)n1 * n2 * n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(id % n3);
        const int iy = (int)((id / n3) % n2);
        const int iz = (int)(id / ((long long)n3 * n2));
        int gz = iz - n1 / 2; if (gz < 0) gz += nfz;
        int gy = iy - n2 / 2; if (gy < 0) gy += nfy;
        int gx = ix - n3 / 2; if (gx < 0) gx += nfx;
        const long long fi = ((long long)gz * nfy + gy) * nfx + gx;
        const float cf = cz[iz] * cy[iy] * cxr[ix];
        if (REAL) {
            float *outs[3] = {o0, o1, o2};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float v = fine[fi * NC + c].x * cf * scale;
                outs[c][id] = accumulate ? outs[c][id] + v : v;
            }
        } else {
            cout[id] = cscale(fine[fi * NC], cf * scale);
        }
    }
}

// ------------------------------------------------------------------ interpolation (type 2)
// HALF: NC = 3 real-volume components combined with the beam direction,
// s = sum_i d_i c_i(q), written to the point and conj(s) to its mirror.
// !HALF: NC = 1, every point.
template <int W, bool HALF>
__global__ __launch_bounds__(256) void k_interp(const float2 *__restrict__ fine, float2 *__restrict__ c,
                                                const Geo *__restrict__ geo, long long P, int R, int C, int nfz,
                                                int nfy, int nfx, float beta, PointFrame fr) {
    constexpr int NC = HALF ? 3 : 1;
    const long long S = (long long)R * C;
    const long long H = HALF ? (S + 1) / 2 : S;
    const long long total = P * H;
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * blockDim.x) {
        const long long p = gid / H;
        const long long idx = gid - p * H;
        const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
        const Geo g = geo[p];
        double tx, ty, tz;
        point_coords(g, fr, j, k, tx, ty, tz);
        int ix, iy, iz;
        float wx[W], wy[W], wz[W];
        es_weights<W>(tx, nfx, beta, ix, wx);
        es_weights<W>(ty, nfy, beta, iy, wy);
        es_weights<W>(tz, nfz, beta, iz, wz);
        float2 acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = make_float2(0.f, 0.f);
        int xs[W];
#pragma unroll
        for (int e = 0; e < W; ++e) xs[e] = wrapi(ix + e, nfx);
        for (int a = 0; a < W; ++a) {
            const int zz = wrapi(iz + a, nfz);
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const int yy = wrapi(iy + b, nfy);
                const float2 *row = fine + ((long long)zz * nfy + yy) * (long long)nfx * NC;
                float2 racc[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) racc[q] = make_float2(0.f, 0.f);
#pragma unroll
                for (int e = 0; e < W; ++e) {
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        const float2 v = row[(long long)xs[e] * NC + q];
                        racc[q].x += v.x * wx[e];
                        racc[q].y += v.y * wx[e];
                    }
                }
                const float wzy = wz[a] * wy[b];
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    acc[q].x += racc[q].x * wzy;
                    acc[q].y += racc[q].y * wzy;
                }
            }
        }
        if constexpr (HALF) {
            const float d0 = (float)g.d0, d1 = (float)g.d1, d2 = (float)g.d2;
            float2 s;
            s.x = d0 * acc[0].x + d1 * acc[1].x + d2 * acc[2].x;
            s.y = d0 * acc[0].y + d1 * acc[1].y + d2 * acc[2].y;
            c[p * S + idx] = s;
            const long long mir = S - 1 - idx;
            if (mir != idx) c[p * S + mir] = cconj(s);
        } else {
            c[p * S + idx] = acc[0];
        }
    }
}

// ------------------------------------------------------------------ spreading (type 1)
template <int W, bool HALF>
__global__ __launch_bounds__(256) void k_spread_atomic(float2 *__restrict__ fine, const float2 *__restrict__ c,
                                                       const Geo *__restrict__ geo, long long P, int R, int C,
                                                       int nfz, int nfy, int nfx, float beta, PointFrame fr) {
    constexpr int NC = HALF ? 3 : 1;
    const long long S = (long long)R * C;
    const long long H = HALF ? (S + 1) / 2 : S;
    const long long total = P * H;
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * blockDim.x) {
        const long long p = gid / H;
        const long long idx = gid - p * H;
        const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
        const Geo g = geo[p];
        float2 v = c[p * S + idx];
        if (HALF) {
            const long long mir = S - 1 - idx;
            if (mir != idx) v = cadd(v, cconj(c[p * S + mir]));
        }
        float2 vc[NC];
        if constexpr (HALF) {
            vc[0] = cscale(v, (float)g.d0);
            vc[1] = cscale(v, (float)g.d1);
            vc[2] = cscale(v, (float)g.d2);
        } else {
            vc[0] = v;
        }
        double tx, ty, tz;
        point_coords(g, fr, j, k, tx, ty, tz);
        int ix, iy, iz;
        float wx[W], wy[W], wz[W];
        es_weights<W>(tx, nfx, beta, ix, wx);
        es_weights<W>(ty, nfy, beta, iy, wy);
        es_weights<W>(tz, nfz, beta, iz, wz);
        for (int a = 0; a < W; ++a) {
            const int zz = wrapi(iz + a, nfz);
            for (int b = 0; b < W; ++b) {
                const int yy = wrapi(iy + b, nfy);
                float2 *row = fine + ((long long)zz * nfy + yy) * (long long)nfx * NC;
                const float wzy = wz[a] * wy[b];
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    const int xx = wrapi(ix + e, nfx);
                    const float wgt = wzy * wx[e];
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        float *dst = reinterpret_cast<float *>(row + (long long)xx * NC + q);
                        atomicAdd(dst, vc[q].x * wgt);
                        atomicAdd(dst + 1, vc[q].y * wgt);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------ host helpers
size_t next_smooth_even(size_t n) {
    for (;; ++n) {
        if (n % 2) continue;
        size_t r = n;
        for (size_t p : {2, 3, 5}) while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

void gauss_legendre(int n, double *x, double *wt) {
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = z;
        wt[i] = 2.0 / ((1 - z * z) * pp * pp);
    }
}

PointFrame frame(const Grid &g, const NufftPlan &pl) {
    PointFrame f;
    f.dX = kTwoPi / (double)g.C;
    f.dY = kTwoPi / (double)g.R;
    f.k0 = 0.5 * (double)(g.C - 1);
    f.j0 = 0.5 * (double)(g.R - 1);
    f.sx = (double)pl.nf[2] / kTwoPi;
    f.sy = (double)pl.nf[1] / kTwoPi;
    f.sz = (double)pl.nf[0] / kTwoPi;
    return f;
}

#define LAM_WIDTH_SWITCH(w, CALL)                                          \
    switch (w) {                                                           \
        case 4: { constexpr int W = 4; CALL; } break;                      \
        case 6: { constexpr int W = 6; CALL; } break;                      \
        case 8: { constexpr int W = 8; CALL; } break;                      \
        case 10: { constexpr int W = 10; CALL; } break;                    \
        case 12: { constexpr int W = 12; CALL; } break;                    \
        default: throw Error(LAM_ERR_NUFFT, "Error in finufft_makeplan: unsupported kernel width"); \
    }

void launch_interp(Ctx &cx, Grid &g, NufftPlan &pl, float2 *c, bool half) {
    const PointFrame fr = frame(g, pl);
    const long long S = g.R * g.C;
    const long long total = g.P * (half ? (S + 1) / 2 : S);
    const unsigned grid = grid_for((size_t)total, 256, 1u << 20);
    const Geo *geo = g.geo.as<Geo>();
    if (half) {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_interp<W, true>), dim3(grid), dim3(256), 0, cx.stream,
                                                  pl.fine.as<float2>(), c, geo, g.P, (int)g.R, (int)g.C, pl.nf[0],
                                                  pl.nf[1], pl.nf[2], pl.beta, fr));
    } else {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_interp<W, false>), dim3(grid), dim3(256), 0, cx.stream,
                                                  pl.fine.as<float2>(), c, geo, g.P, (int)g.R, (int)g.C, pl.nf[0],
                                                  pl.nf[1], pl.nf[2], pl.beta, fr));
    }
    check_launch();
}

void launch_spread(Ctx &cx, Grid &g, NufftPlan &pl, const float2 *c, bool half) {
    const PointFrame fr = frame(g, pl);
    const long long S = g.R * g.C;
    const long long total = g.P * (half ? (S + 1) / 2 : S);
    const unsigned grid = grid_for((size_t)total, 256, 1u << 20);
    const Geo *geo = g.geo.as<Geo>();
    const int nc = half ? 3 : 1;
    LAM_HIP(hipMemsetAsync(pl.fine.ptr, 0, sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * nc, cx.stream));
    if (half) {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_spread_atomic<W, true>), dim3(grid), dim3(256), 0, cx.stream,
                                                  pl.fine.as<float2>(), c, geo, g.P, (int)g.R, (int)g.C, pl.nf[0],
                                                  pl.nf[1], pl.nf[2], pl.beta, fr));
    } else {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_spread_atomic<W, false>), dim3(grid), dim3(256), 0, cx.stream,
                                                  pl.fine.as<float2>(), c, geo, g.P, (int)g.R, (int)g.C, pl.nf[0],
                                                  pl.nf[1], pl.nf[2], pl.beta, fr));
    }
    check_launch();
}

void ensure_fine(NufftPlan &pl, int nc) {
    pl.fine.ensure(sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * nc);
}

void deapod_pad(Ctx &cx, NufftPlan &pl, const float *const m[3], const float2 *cin, int nc) {
    const long long total = (long long)pl.nf[0] * pl.nf[1] * pl.nf[2];
    const unsigned grid = grid_for((size_t)total, 256, 1u << 16);
    if (nc == 3)
        hipLaunchKernelGGL((k_deapod_pad<3, true>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(), m[0],
                           m[1], m[2], nullptr, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0], pl.nf[1],
                           pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(), pl.corr[2].as<float>());
    else
        hipLaunchKernelGGL((k_deapod_pad<1, false>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           nullptr, nullptr, nullptr, cin, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>());
    check_launch();
}

void crop_deapod(Ctx &cx, NufftPlan &pl, float *const out[3], float2 *cout, int nc, float scale, bool accumulate) {
    const long long total = pl.n[0] * pl.n[1] * pl.n[2];
    const unsigned grid = grid_for((size_t)total, 256, 1u << 16);
    if (nc == 3)
        hipLaunchKernelGGL((k_crop_deapod<3, true>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           out[0], out[1], out[2], nullptr, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>(), scale, accumulate ? 1 : 0);
    else
        hipLaunchKernelGGL((k_crop_deapod<1, false>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           nullptr, nullptr, nullptr, cout, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>(), scale, 0);
    check_launch();
}

void check_dims(const long long dims[3]) {
    for (int a = 0; a < 3; ++a)
        LAM_REQUIRE(dims[a] >= 1 && dims[a] <= (1 << 20), LAM_ERR_ARG, "invalid volume dimensions");
}

}  // namespace

Grid::~Grid() = default;

int width_for_eps(double eps) {
    if (!(eps > 0)) eps = 1e-5;
    int w = (int)std::ceil(-std::log10(eps / 10.0));
    if (w % 2) ++w;
    return std::max(4, std::min(12, w));
}

NufftPlan &get_nufft_plan(Ctx &cx, const long long dims[3], double eps) {
    check_dims(dims);
    const int w = width_for_eps(eps);
    auto key = std::make_tuple(dims[0], dims[1], dims[2] * 64 + w, 0);
    auto it = cx.nufft_plans.find(key);
    if (it != cx.nufft_plans.end()) return *it->second;
    auto pl = std::make_unique<NufftPlan>();
    pl->w = w;
    pl->beta = (float)(2.30 * w);
    const int nq = 200;
    std::vector<double> xq(nq), wq(nq);
    gauss_legendre(nq, xq.data(), wq.data());
    const double pi = 3.14159265358979323846;
    for (int a = 0; a < 3; ++a) {
        pl->n[a] = dims[a];
        size_t f = next_smooth_even(std::max<size_t>(2 * (size_t)dims[a], 2 * (size_t)w));
        pl->nf[a] = (int)f;
        std::vector<float> corr(dims[a]);
        for (long long i = 0; i < dims[a]; ++i) {
            const long long k = i - dims[a] / 2;
            double s = 0;
            for (int q = 0; q < nq; ++q) {
                const double z = xq[q];
                const double phi = std::exp((double)pl->beta * (std::sqrt(1.0 - z * z) - 1.0));
                s += wq[q] * phi * std::cos(pi * (double)k * w * z / (double)f);
            }
            corr[i] = (float)(1.0 / (0.5 * w * s));
        }
        pl->corr[a].alloc(sizeof(float) * dims[a]);
        LAM_HIP(hipMemcpy(pl->corr[a].ptr, corr.data(), sizeof(float) * dims[a], hipMemcpyHostToDevice));
    }
    NufftPlan &ref = *pl;
    cx.nufft_plans.emplace(key, std::move(pl));
    return ref;
}

std::unique_ptr<Grid> make_grid(Ctx &cx, const double *alpha, const double *gamma, long long P, long long R,
                                long long C) {
    LAM_REQUIRE(P >= 1 && R >= 1 && C >= 1, LAM_ERR_ARG, "invalid polar grid shape");
    LAM_REQUIRE(R * C < (1ll << 40), LAM_ERR_ARG, "polar grid too large");
    auto g = std::make_unique<Grid>();
    g->cx = &cx;
    g->P = P; g->R = R; g->C = C;
    g->alpha.assign(alpha, alpha + P);
    g->gamma.assign(gamma, gamma + P);
    g->geo_host.resize(P);
    for (long long p = 0; p < P; ++p) {
        const double a = alpha[p], gm = gamma[p];
        Geo e;
        e.ux = std::cos(gm);
        e.uz = -std::sin(gm);
        e.vx = std::cos(a) * std::sin(gm);
        e.vy = -std::sin(a);
        e.vz = std::cos(a) * std::cos(gm);
        e.d0 = std::sin(gm) * std::sin(a);
        e.d1 = std::cos(a);
        e.d2 = std::cos(gm) * std::sin(a);
        g->geo_host[p] = e;
    }
    g->geo.alloc(sizeof(Geo) * P);
    LAM_HIP(hipMemcpy(g->geo.ptr, g->geo_host.data(), sizeof(Geo) * P, hipMemcpyHostToDevice));
    return g;
}

void nufft2_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, float2 *c, const float2 *f) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 1);
    deapod_pad(cx, pl, nullptr, f, 1);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 1, -1);
    launch_interp(cx, g, pl, c, false);
}

void nufft1_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, const float2 *c, float2 *f) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 1);
    launch_spread(cx, g, pl, c, false);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 1, +1);
    crop_deapod(cx, pl, nullptr, f, 1, 1.0f, false);
}

void forward(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const m[3], float *proj) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    deapod_pad(cx, pl, m, nullptr, 3);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 3, -1);
    launch_interp(cx, g, pl, cbuf, true);
    det_fft2(cx, cbuf, false, proj, true, (int)g.P, (int)g.R, (int)g.C, +1, 1.0f / (float)(g.R * g.C), cbuf);
}

void adjoint(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *proj, float *const m[3],
             bool accumulate) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    det_fft2(cx, proj, true, cbuf, false, (int)g.P, (int)g.R, (int)g.C, -1, 1.0f, cbuf);
    launch_spread(cx, g, pl, cbuf, true);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 3, +1);
    crop_deapod(cx, pl, m, nullptr, 3, 1.0f / (float)(g.R * g.C), accumulate);
}

void sysmat_nufft(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const x[3],
                  float *const y[3]) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    deapod_pad(cx, pl, x, nullptr, 3);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 3, -1);
    launch_interp(cx, g, pl, cbuf, true);
    launch_spread(cx, g, pl, cbuf, true);
    fft3d(cx, pl.fine.as<float2>(), pl.nf[0], pl.nf[1], pl.nf[2], 3, +1);
    crop_deapod(cx, pl, y, nullptr, 3, 1.0f / (float)(g.R * g.C), false);
}

}  // namespace lam
