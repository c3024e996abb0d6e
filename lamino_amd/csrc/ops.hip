// Fused elementwise, stencil, reduction and preconditioner kernels.
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "fft.h"
#include "ops.h"

namespace lam {

namespace {

struct RedSpec {
    RedTerm t[kMaxRedTerms];
    int nterms;
    long long n;
};

__global__ __launch_bounds__(256) void k_reduce(RedSpec s, double *partials) {
    double acc[kMaxRedTerms];
#pragma unroll
    for (int q = 0; q < kMaxRedTerms; ++q) acc[q] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < s.n; i += stride) {
#pragma unroll
        for (int q = 0; q < kMaxRedTerms; ++q) {
            if (q < s.nterms) {
                const RedTerm &t = s.t[q];
                const double a = t.a[i];
                if (t.kind == RED_DOT) acc[q] += a * (double)t.b[i];
                else if (t.kind == RED_DIFF2) { const double d = a - (double)t.b[i]; acc[q] += d * d; }
                else acc[q] += a * a;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kMaxRedTerms; ++q) {
        if (q < s.nterms) {
            const double v = block_sum_256(acc[q]);
            if (threadIdx.x == 0) partials[(size_t)q * gridDim.x + blockIdx.x] = v;
        }
    }
}

// The same sums for exactly NT terms (NT <= 4, the solvers' usual counts), four grid
// strides per iteration: all their loads are issued before the first add (one load
// pair in flight per thread ran at ~1.9 TB/s), and the adds keep the element order
// i, i + stride, ..., so every partial is the same bit for bit as k_reduce's.
template <int NT>
__global__ __launch_bounds__(256) void k_reduce_nt(RedSpec s, double *partials) {
    double acc[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[q] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    auto add = [&](int q, float av, float bv) {
        const int kind = s.t[q].kind;
        const double a = av;
        if (kind == RED_DOT) acc[q] += a * (double)bv;
        else if (kind == RED_DIFF2) { const double d = a - (double)bv; acc[q] += d * d; }
        else acc[q] += a * a;
    };
    for (; i + 3 * stride < s.n; i += 4 * stride) {
        float a[NT][4], b[NT][4];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const RedTerm &t = s.t[q];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[q][u] = t.a[i + u * stride];
#pragma unroll
            for (int u = 0; u < 4; ++u) b[q][u] = t.kind != RED_SQ ? t.b[i + u * stride] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < NT; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) add(q, a[q][u], b[q][u]);
    }
    for (; i < s.n; i += stride)
#pragma unroll
        for (int q = 0; q < NT; ++q) add(q, s.t[q].a[i], s.t[q].kind != RED_SQ ? s.t[q].b[i] : 0.f);
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const double v = block_sum_256(acc[q]);
        if (threadIdx.x == 0) partials[(size_t)q * gridDim.x + blockIdx.x] = v;
    }
}

// fixed-order final sum of the partials of each term
__global__ __launch_bounds__(256) void k_reduce_final(const double *partials, int nblocks, int nterms, double *out) {
    for (int q = 0; q < nterms; ++q) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[(size_t)q * nblocks + b];
        const double v = block_sum_256(acc);
        if (threadIdx.x == 0) out[q] = v;
    }
}

__global__ __launch_bounds__(256) void k_max_partial(const float *a, long long n, float *partials) {
    __shared__ float sh[256];
    float m = -INFINITY;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, a[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// ------------------------------------------------------------------ stencils
__device__ __forceinline__ bool interior(int i, int j, int k, int n1, int n2, int n3) {
    return i >= 1 && i < n1 - 1 && j >= 1 && j < n2 - 1 && k >= 1 && k < n3 - 1;
}

__global__ __launch_bounds__(256) void k_grad_u(const float *__restrict__ u, float *g0, float *g1, float *g2, int n1,
                                                int n2, int n3, Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s1 = (long long)n2 * n3, s2 = n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        float a = 0.f, b = 0.f, c = 0.f;
        if (interior(i, j, k, n1, n2, n3)) {
            a = (u[id + s1] - u[id - s1]) * 0.5f;
            b = (u[id + s2] - u[id - s2]) * 0.5f;
            c = (u[id + 1] - u[id - 1]) * 0.5f;
        }
        g0[id] = a; g1[id] = b; g2[id] = c;
    }
}

__global__ __launch_bounds__(256) void k_divergence(const float *__restrict__ v0, const float *__restrict__ v1,
                                                    const float *__restrict__ v2, float *out, int n1, int n2, int n3,
                                                    Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s1 = (long long)n2 * n3, s2 = n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        float r = 0.f;
        if (interior(i, j, k, n1, n2, n3)) {
            const float A = v0[id + s1] - v0[id - s1];
            const float B = v1[id + s2] - v1[id - s2];
            const float C = v2[id + 1] - v2[id - 1];
            r = (float)(-(double)A / 2.0 - (double)B / 2.0 - (double)C / 2.0);
        }
        out[id] = r;
    }
}

// grad_u component along an axis at a point that is displaced by +-1 along it
// from an interior voxel: nonzero only when that point is itself interior.
__device__ __forceinline__ float lap_at(const float *u, long long id, int i, int j, int k, int n1, int n2, int n3,
                                        long long s1, long long s2) {
    // axis 0 (n1): g0(i+1) - g0(i-1)
    const float g0p = (i + 1 < n1 - 1) ? (u[id + 2 * s1] - u[id]) * 0.5f : 0.f;
    const float g0m = (i - 1 >= 1) ? (u[id] - u[id - 2 * s1]) * 0.5f : 0.f;
    const float g1p = (j + 1 < n2 - 1) ? (u[id + 2 * s2] - u[id]) * 0.5f : 0.f;
    const float g1m = (j - 1 >= 1) ? (u[id] - u[id - 2 * s2]) * 0.5f : 0.f;
    const float g2p = (k + 1 < n3 - 1) ? (u[id + 2] - u[id]) * 0.5f : 0.f;
    const float g2m = (k - 1 >= 1) ? (u[id] - u[id - 2]) * 0.5f : 0.f;
    const float A = g0p - g0m, B = g1p - g1m, C = g2p - g2m;
    return (float)(-(double)A / 2.0 - (double)B / 2.0 - (double)C / 2.0);
}

__global__ __launch_bounds__(256) void k_laplacian(const float *__restrict__ u, float *out, int n1, int n2, int n3,
                                                   Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s1 = (long long)n2 * n3, s2 = n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        out[id] = interior(i, j, k, n1, n2, n3) ? lap_at(u, id, i, j, k, n1, n2, n3, s1, s2) : 0.f;
    }
}

__global__ __launch_bounds__(256) void k_add_mu_lap3(CVol3 x, Vol3 y, float mu, int n1, int n2, int n3, Split3 sp) {
    const long long total = (long long)n1 * n2 * n3;
    const long long s1 = (long long)n2 * n3, s2 = n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(id, sp, n2, n3, i, j, k);
        if (!interior(i, j, k, n1, n2, n3)) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) y.c[c][id] += lap_at(x.c[c], id, i, j, k, n1, n2, n3, s1, s2) * mu;
    }
}

// ------------------------------------------------------------------ qGGMRF
__constant__ float c_qw[27] = {0.0302f, 0.037f, 0.0302f, 0.037f, 0.0523f, 0.037f, 0.0302f, 0.037f, 0.0302f,
                               0.037f,  0.0523f, 0.037f, 0.0532f, 0.f,    0.0523f, 0.037f, 0.0523f, 0.037f,
                               0.0302f, 0.037f, 0.0302f, 0.037f, 0.0523f, 0.037f, 0.0302f, 0.037f, 0.0302f};

// d_potential_fcn with the reference's mixed float/double arithmetic
// (src/qggmrf.cpp:41-57): pow in double, the rest as written there.
__device__ __forceinline__ float dpot(float delta, float sigma_q, float sigma_qp, float p) {
    const double Q = 2.0;
    const float C = 0.001f;
    const float ad = fabsf(delta);
    if (ad > 0.f) {
        const float t1 = (float)(pow((double)ad, Q - (double)p) / (double)sigma_qp);
        const float t3 = C + t1;
        return (float)((double)(delta / (t3 * sigma_q)) * (Q - ((Q - (double)p) * (double)t1) / (double)t3));
    }
    return (float)(Q / ((double)sigma_q * 0.001));
}

// The prior's 26 terms, each edge evaluated once. For δ ≠ 0, dpot is odd bit for bit:
// |δ| feeds the power, the sign enters only through δ / (t3·σ_q), and negating a float
// or double quotient or product is exact; so with v = u + o the term of the edge seen
// from v is the negated term seen from u (x_v − x_u = −(x_u − x_v) exactly). For δ = 0
// (and NaN) dpot is the constant 2/(σ_q·0.001) from both ends. A workgroup takes a
// (j, k) tile of kQJ−2 × kQK−2 interior voxels with a one-voxel halo and marches it
// along i over kQPlanes planes. On each plane every voxel of the extended tile
// evaluates its 13 forward edges (o > 0 in (a, b, c) order: 9 to plane i+1, 4 in the
// plane), keeps them for its own sum and puts the value its neighbour needs (negated
// unless δ = 0) in LDS; each interior voxel then adds its 26 terms in
// src/qggmrf.cpp:74-90's (a, b, c) order, the 13 backward ones from LDS (plane i−1's
// edges to plane i, the plane's own in-plane edges). The sums are those of a per-voxel
// loop over the 26 neighbours, bit for bit, with 13/26 of the FP64 powers plus the
// halo's (0.61 of them for a full tile).
constexpr int kQJ = 16, kQK = 32, kQPlanes = 16;

__global__ __launch_bounds__(kQJ * kQK) void k_qggmrf(Vol3 gv, CVol3 xv, int n1, int n2, int n3, float sigma_q,
                                                    float sigma_qp, float p, int tiles_k) {
    // one component per blockIdx.z: the three share one launch (no half-empty last round per component)
    const unsigned comp = blockIdx.z;
    float *g = comp == 0 ? gv.c[0] : comp == 1 ? gv.c[1] : gv.c[2];
    const float *__restrict__ x = comp == 0 ? xv.c[0] : comp == 1 ? xv.c[1] : xv.c[2];
    // Backward values: plane i−1's edges (1, b, c), slot (b+1)·3 + c+1, and plane i's
    // edges (0,0,1), (0,1,−1), (0,1,0), (0,1,1), each as seen from the edge's far end.
    __shared__ float up[9][kQJ * kQK];
    __shared__ float inpl[4][kQJ * kQK];
    const int t = threadIdx.x, jj = t / kQK, kk = t % kQK;
    const int j = (blockIdx.x / tiles_k) * (kQJ - 2) + jj;  // the halo row/column is jj, kk = 0
    const int k = (blockIdx.x % tiles_k) * (kQK - 2) + kk;
    const int i_lo = 1 + blockIdx.y * kQPlanes, i_hi = min(i_lo + kQPlanes, n1 - 1);
    const bool inb = j < n2 && k < n3;
    const bool own = jj >= 1 && jj < kQJ - 1 && kk >= 1 && kk < kQK - 1 && j < n2 - 1 && k < n3 - 1;
    const long long s1 = (long long)n2 * n3;
    // Edge values of voxel (i, j, k); an edge leaving the volume is 0 (no interior voxel
    // reads it). Bit s of `flat` is set where δ is 0 or NaN.
    auto edge = [&](float d, unsigned &flat, int s) {
        if (!(fabsf(d) > 0.f)) flat |= 1u << s;
        return dpot(d, sigma_q, sigma_qp, p);
    };
    auto back = [](float e, unsigned flat, int s) { return (flat >> s) & 1u ? e : -e; };
    auto up_edges = [&](int i, float e[9], unsigned &flat) {
        const long long id = i * s1 + (long long)j * n3 + k;
        const float u = inb ? x[id] : 0.f;
#pragma unroll
        for (int b = -1; b < 2; ++b)
#pragma unroll
            for (int c = -1; c < 2; ++c) {
                float v = 0.f;
                if (inb && j + b >= 0 && j + b < n2 && k + c >= 0 && k + c < n3)
                    v = edge(x[id + s1 + b * n3 + c] - u, flat, (b + 1) * 3 + (c + 1));
                e[(b + 1) * 3 + (c + 1)] = v;
            }
    };
    {
        float e[9];
        unsigned flat = 0;
        up_edges(i_lo - 1, e, flat);
#pragma unroll
        for (int s = 0; s < 9; ++s) up[s][t] = back(e[s], flat, s);
    }
    for (int i = i_lo; i < i_hi; ++i) {
        float eu[9], ep[4] = {0.f, 0.f, 0.f, 0.f};
        unsigned flat = 0;  // bits 0-8: eu, 9-12: ep
        up_edges(i, eu, flat);
        const long long id = i * s1 + (long long)j * n3 + k;
        if (inb) {
            const float u = x[id];
            if (k + 1 < n3) ep[0] = edge(x[id + 1] - u, flat, 9);
            if (j + 1 < n2) {
#pragma unroll
                for (int c = -1; c < 2; ++c)
                    if (k + c >= 0 && k + c < n3) ep[c + 2] = edge(x[id + n3 + c] - u, flat, 11 + c);
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) inpl[s][t] = back(ep[s], flat, 9 + s);
        __syncthreads();
        if (own) {
            float sum = 0.f;
#pragma unroll
            for (int b = -1; b < 2; ++b)
#pragma unroll
                for (int c = -1; c < 2; ++c)
                    sum += c_qw[(b + 1) * 3 + (c + 1)] * up[(1 - b) * 3 + (1 - c)][t + b * kQK + c];
            sum += c_qw[9] * inpl[3][t - kQK - 1];
            sum += c_qw[10] * inpl[2][t - kQK];
            sum += c_qw[11] * inpl[1][t - kQK + 1];
            sum += c_qw[12] * inpl[0][t - 1];
            sum += c_qw[14] * ep[0];
            sum += c_qw[15] * ep[1];
            sum += c_qw[16] * ep[2];
            sum += c_qw[17] * ep[3];
#pragma unroll
            for (int s = 0; s < 9; ++s) sum += c_qw[18 + s] * eu[s];
            g[id] += sum;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 9; ++s) up[s][t] = back(eu[s], flat, s);
    }
}

// ------------------------------------------------------------------ preconditioner filter
// Full-spectrum ramp |k| of the half-spectrum filter of include/precond.h:41-61
// (the c2r Hermitian extension makes it even in both axes).
__global__ __launch_bounds__(256) void k_ramp_filter(float2 *a, float2 *b, int n1, int n2, int n3) {
    const long long total = (long long)n1 * n2 * n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(id % n3), j = (int)((id / n3) % n2);
        const int kh = k <= n3 / 2 ? k : n3 - k;
        const float xf = (float)kh / (float)n3;
        const float yf = (j <= n2 / 2) ? (float)j / (float)n2 : (float)(j - n2) / (float)n2;
        const float f = sqrtf(xf * xf + yf * yf);
        a[id] = cscale(a[id], f);
        if (b) b[id] = cscale(b[id], f);
    }
}

// ------------------------------------------------------------------ elementwise
__global__ __launch_bounds__(256) void k_axpby3(long long n, float a, CVol3 x, float b, Vol3 y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (y.c[c]) y.c[c][i] = a * x.c[c][i] + b * y.c[c][i];
    }
}

__global__ __launch_bounds__(256) void k_copy3(long long n, CVol3 x, Vol3 y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (y.c[c]) y.c[c][i] = x.c[c][i];
    }
}

__global__ __launch_bounds__(256) void k_fill(long long n, float *y, float v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = v;
}

__global__ __launch_bounds__(256) void k_scale_to(long long n, const float *x, float s, float *y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i] * s;
}

unsigned vgrid(long long n) { return grid_for((size_t)n, 256, 1u << 16); }

}  // namespace

void reduce_terms_dev(Ctx &cx, long long n, const RedTerm *terms, int nterms, double *out_dev, double *partials) {
    LAM_REQUIRE(nterms >= 1 && nterms <= kMaxRedTerms, LAM_ERR_ARG, "bad reduction");
    RedSpec s;
    for (int q = 0; q < nterms; ++q) s.t[q] = terms[q];
    for (int q = nterms; q < kMaxRedTerms; ++q) s.t[q] = RedTerm{RED_SQ, nullptr, nullptr};
    s.nterms = nterms;
    s.n = n;
    if (!partials) {
        cx.red.ensure(sizeof(double) * kRedPartials);
        partials = cx.red.as<double>();
    }
    switch (nterms) {
        case 1: hipLaunchKernelGGL(k_reduce_nt<1>, dim3(kRedBlocks), dim3(256), 0, cx.stream, s, partials); break;
        case 2: hipLaunchKernelGGL(k_reduce_nt<2>, dim3(kRedBlocks), dim3(256), 0, cx.stream, s, partials); break;
        case 3: hipLaunchKernelGGL(k_reduce_nt<3>, dim3(kRedBlocks), dim3(256), 0, cx.stream, s, partials); break;
        case 4: hipLaunchKernelGGL(k_reduce_nt<4>, dim3(kRedBlocks), dim3(256), 0, cx.stream, s, partials); break;
        default: hipLaunchKernelGGL(k_reduce, dim3(kRedBlocks), dim3(256), 0, cx.stream, s, partials); break;
    }
    check_launch();
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, cx.stream, partials, kRedBlocks, nterms, out_dev);
    check_launch();
}

void reduce_terms(Ctx &cx, long long n, const RedTerm *terms, int nterms, double *out_host) {
    cx.scal.ensure(sizeof(double) * 64);
    reduce_terms_dev(cx, n, terms, nterms, cx.scal.as<double>());
    LAM_HIP(hipMemcpyAsync(out_host, cx.scal.ptr, sizeof(double) * nterms, hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
}

// ---- request ingest / output helpers
namespace {
__global__ __launch_bounds__(256) void k_mask_nonfinite(float *a, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!isfinite(a[i])) a[i] = 0.f;
}

// out[j][k][i] = in[i][j][k]: a 32 x 32 tile of the (i, jk) plane through LDS, so
// that both the reads (along jk) and the writes (along i) are coalesced.
__global__ __launch_bounds__(256) void k_transpose120(const float *__restrict__ in, float *__restrict__ out, int n1,
                                                      long long m) {
    __shared__ float t[32][33];
    const long long c0 = (long long)blockIdx.x * 32;   // jk
    const int r0 = (int)blockIdx.y * 32;               // i
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int i = r0 + r;
        const long long jk = c0 + tx;
        if (i < n1 && jk < m) t[r][tx] = in[(long long)i * m + jk];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long jk = c0 + r;
        const int i = r0 + tx;
        if (i < n1 && jk < m) out[jk * n1 + i] = t[tx][r];
    }
}
}  // namespace

namespace {
// b[k + o1][i + o2][j + o3] = a[i][j][k]: array::transpose(a, {2, 0, 1})
// (include/array_ops.h:117-137) placed into a symmetric zero padding (pad3d,
// include/padding.h:168-204), as src/forward.cpp:123-140 prepares a sample.
__global__ __launch_bounds__(256) void k_transpose201_pad(const float *__restrict__ a, float *__restrict__ b, int d0,
                                                          int d1, int d2, int p1, int p2, int o1, int o2, int o3) {
    const long long total = (long long)d0 * d1 * d2;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(id % d2), j = (int)((id / d2) % d1), i = (int)(id / ((long long)d2 * d1));
        b[(((long long)(k + o1)) * p1 + (i + o2)) * p2 + (j + o3)] = a[id];
    }
}

// out[p][r][c] = in[p][r + o1][c + o2]  (crop2d SYMMETRIC, include/padding.h:134-165)
__global__ __launch_bounds__(256) void k_crop2(const float *__restrict__ in, float *__restrict__ out, long long P,
                                               int R, int C, int R2, int C2, int o1, int o2) {
    const long long total = P * R * C;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(id % C);
        const long long pr = id / C;
        const int r = (int)(pr % R);
        const long long p = pr / R;
        out[id] = in[(p * R2 + r + o1) * C2 + c + o2];
    }
}

// Demagnetising field in Fourier space (include/demag.h:13-79): with k the angular
// frequencies 2 pi f / n of each axis (f in numpy's fftfreq order),
// H^(k) = -k (k . m^(k)) / |k|^2, H^(0) = 0.  s holds [n1][n2][n3][3] spectra.
__global__ __launch_bounds__(256) void k_demag(float2 *__restrict__ s, int n1, int n2, int n3) {
    const long long total = (long long)n1 * n2 * n3;
    const double d1 = 2.0 * 3.14159265358979323846 / n1, d2 = 2.0 * 3.14159265358979323846 / n2,
                 d3 = 2.0 * 3.14159265358979323846 / n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(id % n3), b = (int)((id / n3) % n2), a = (int)(id / ((long long)n3 * n2));
        const float kx = (float)((a < (n1 + 1) / 2 ? a : a - n1) * d1);
        const float ky = (float)((b < (n2 + 1) / 2 ? b : b - n2) * d2);
        const float kz = (float)((c < (n3 + 1) / 2 ? c : c - n3) * d3);
        float2 *v = s + 3 * id;
        if (id == 0) {
            v[0] = v[1] = v[2] = make_float2(0.f, 0.f);
            continue;
        }
        const float kinv = 1.0f / (kx * kx + ky * ky + kz * kz);
        const float2 km = make_float2(kx * v[0].x + ky * v[1].x + kz * v[2].x, kx * v[0].y + ky * v[1].y + kz * v[2].y);
        v[0] = make_float2(-km.x * kx * kinv, -km.y * kx * kinv);
        v[1] = make_float2(-km.x * ky * kinv, -km.y * ky * kinv);
        v[2] = make_float2(-km.x * kz * kinv, -km.y * kz * kinv);
    }
}

__global__ __launch_bounds__(256) void k_interleave3(const float *__restrict__ m0, const float *__restrict__ m1,
                                                     const float *__restrict__ m2, float2 *__restrict__ s, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        s[3 * i] = make_float2(m0[i], 0.f);
        s[3 * i + 1] = make_float2(m1[i], 0.f);
        s[3 * i + 2] = make_float2(m2[i], 0.f);
    }
}

__global__ __launch_bounds__(256) void k_deinterleave3(const float2 *__restrict__ s, float *__restrict__ h0,
                                                       float *__restrict__ h1, float *__restrict__ h2, long long n,
                                                       float scale) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        h0[i] = s[3 * i].x * scale;
        h1[i] = s[3 * i + 1].x * scale;
        h2[i] = s[3 * i + 2].x * scale;
    }
}
}  // namespace

void transpose201_pad(Ctx &cx, const long long d[3], const float *in, const long long pd[3], float *out) {
    const long long total = d[0] * d[1] * d[2];
    LAM_REQUIRE(pd[0] >= d[2] && pd[1] >= d[0] && pd[2] >= d[1], LAM_ERR_ARG, "padding smaller than the volume");
    hipLaunchKernelGGL(k_transpose201_pad, dim3(grid_for((size_t)total, 256, 1u << 16)), dim3(256), 0, cx.stream, in,
                       out, (int)d[0], (int)d[1], (int)d[2], (int)pd[1], (int)pd[2], (int)((pd[0] - d[2]) / 2),
                       (int)((pd[1] - d[0]) / 2), (int)((pd[2] - d[1]) / 2));
    check_launch();
}

void crop2(Ctx &cx, const float *in, long long P, int R2, int C2, float *out, int R, int C) {
    LAM_REQUIRE(R <= R2 && C <= C2, LAM_ERR_ARG, "crop larger than the projections");
    hipLaunchKernelGGL(k_crop2, dim3(grid_for((size_t)(P * R * C), 256, 1u << 16)), dim3(256), 0, cx.stream, in, out,
                       P, R, C, R2, C2, (R2 - R) / 2, (C2 - C) / 2);
    check_launch();
}

void demag(Ctx &cx, const long long d[3], const float *const m[3], float *const h[3], float2 *work) {
    const long long n = d[0] * d[1] * d[2];
    const unsigned g = grid_for((size_t)n, 256, 1u << 16);
    hipLaunchKernelGGL(k_interleave3, dim3(g), dim3(256), 0, cx.stream, m[0], m[1], m[2], work, n);
    check_launch();
    fft3d(cx, work, (int)d[0], (int)d[1], (int)d[2], 3, -1);
    hipLaunchKernelGGL(k_demag, dim3(g), dim3(256), 0, cx.stream, work, (int)d[0], (int)d[1], (int)d[2]);
    check_launch();
    fft3d(cx, work, (int)d[0], (int)d[1], (int)d[2], 3, +1);
    hipLaunchKernelGGL(k_deinterleave3, dim3(g), dim3(256), 0, cx.stream, work, h[0], h[1], h[2], n,
                       (float)(1.0 / (double)n));
    check_launch();
}

void mask_infs_nans(Ctx &cx, long long n, float *a) {
    hipLaunchKernelGGL(k_mask_nonfinite, dim3(grid_for((size_t)n, 256, 1u << 16)), dim3(256), 0, cx.stream, a, n);
    check_launch();
}

void transpose120_dev(Ctx &cx, const long long d[3], const float *in, float *out) {
    const long long m = d[1] * d[2];
    const long long gx = (m + 31) / 32, gy = (d[0] + 31) / 32;
    LAM_REQUIRE(gx < (1ll << 31) && gy <= 65535, LAM_ERR_ARG, "transpose: volume too large for one launch");
    hipLaunchKernelGGL(k_transpose120, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, cx.stream, in, out, (int)d[0],
                       m);
    check_launch();
}

// ---- the no-allocation dot product (lam_dot_noalloc)
// Its scratch is this file-scope device array: the 1024 FP64 partial sums of the
// fixed reduction grid and the result.  Its size does not depend on n.  Each
// device holds its own copy at its own address until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for the
// device (the array's name is a host-side stand-in, not a device address).
namespace {
constexpr size_t kNoAllocScratchDoubles = (size_t)kRedBlocks + 64;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;   // device -> address of that device's scratch
}  // namespace

// Looked up once per device, with the device current; later calls reuse it.
double *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    LAM_HIP(hipSetDevice(device));
    void *p = nullptr;
    LAM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    LAM_REQUIRE(p != nullptr, LAM_ERR_HIP, "no device address for the no-allocation scratch");
    g_scratch_ptr.emplace(device, static_cast<double *>(p));
    return static_cast<double *>(p);
}

double dot_noalloc(Ctx &cx, long long n, const float *a, const float *b) {
    LAM_REQUIRE(n >= 0 && n <= kNoAllocMaxElems, LAM_ERR_ARG,
                "lam_dot_noalloc: n = " + std::to_string(n) + " is outside [0, 2^40]");
    double *ws = noalloc_scratch(cx.device);
    const RedTerm t{RED_DOT, a, b};
    // the same kernels and launch shape as lam_dot: partials in ws[0, kRedBlocks), the
    // sum in ws[kRedBlocks]; every slot read is written first, so nothing needs zeroing
    reduce_terms_dev(cx, n, &t, 1, ws + kRedBlocks, ws);
    double out = 0.0;
    LAM_HIP(hipMemcpyAsync(&out, ws + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
    return out;
}

float max_value(Ctx &cx, long long n, const float *a) {
    const int nb = 256;
    cx.red.ensure(sizeof(double) * kRedBlocks * kMaxRedTerms);
    hipLaunchKernelGGL(k_max_partial, dim3(nb), dim3(256), 0, cx.stream, a, n, reinterpret_cast<float *>(cx.red.ptr));
    check_launch();
    std::vector<float> h(nb);
    LAM_HIP(hipMemcpyAsync(h.data(), cx.red.ptr, sizeof(float) * nb, hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
    float m = -INFINITY;
    for (float v : h) m = std::fmax(m, v);
    return m;
}

void grad_u(Ctx &cx, const long long d[3], const float *u, float *g0, float *g1, float *g2) {
    const long long n = d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_grad_u, dim3(vgrid(n)), dim3(256), 0, cx.stream, u, g0, g1, g2, (int)d[0], (int)d[1],
                       (int)d[2], make_split3(d));
    check_launch();
}

void divergence(Ctx &cx, const long long d[3], const float *v0, const float *v1, const float *v2, float *out) {
    const long long n = d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_divergence, dim3(vgrid(n)), dim3(256), 0, cx.stream, v0, v1, v2, out, (int)d[0], (int)d[1],
                       (int)d[2], make_split3(d));
    check_launch();
}

void laplacian(Ctx &cx, const long long d[3], const float *u, float *out) {
    const long long n = d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_laplacian, dim3(vgrid(n)), dim3(256), 0, cx.stream, u, out, (int)d[0], (int)d[1], (int)d[2],
                       make_split3(d));
    check_launch();
}

void add_mu_laplacian3(Ctx &cx, const long long d[3], CVol3 x, Vol3 y, float mu) {
    const long long n = d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_add_mu_lap3, dim3(vgrid(n)), dim3(256), 0, cx.stream, x, y, mu, (int)d[0], (int)d[1],
                       (int)d[2], make_split3(d));
    check_launch();
}

void qggmrf(Ctx &cx, const long long d[3], float *g, const float *x, float sigma, float p) {
    qggmrf3(cx, d, Vol3{{g, nullptr, nullptr}}, CVol3{{x, nullptr, nullptr}}, 1, sigma, p);
}

void qggmrf3(Ctx &cx, const long long d[3], Vol3 g, CVol3 x, int ncomp, float sigma, float p) {
    const long long n = d[0] * d[1] * d[2];
    const float sq = (float)std::pow((double)sigma, 2.0);
    const float sqp = (float)std::pow((double)sigma, 2.0 - (double)p);
    if (d[0] < 3 || d[1] < 3 || d[2] < 3 || n == 0) return;  // no interior voxel
    const long long tiles_j = (d[1] - 2 + kQJ - 3) / (kQJ - 2), tiles_k = (d[2] - 2 + kQK - 3) / (kQK - 2);
    const long long chunks = (d[0] - 2 + kQPlanes - 1) / kQPlanes;
    if (tiles_j * tiles_k > 0x7fffffffLL || chunks > 65535)
        throw Error(LAM_ERR_ARG, "qggmrf: volume too large for the tiled prior kernel");
    if (ncomp < 1 || ncomp > 3) throw Error(LAM_ERR_ARG, "qggmrf: 1 to 3 components");
    for (int c = 0; c < ncomp; ++c)
        if (!g.c[c] || !x.c[c]) throw Error(LAM_ERR_ARG, "qggmrf: null component");
    hipLaunchKernelGGL(k_qggmrf, dim3((unsigned)(tiles_j * tiles_k), (unsigned)chunks, (unsigned)ncomp),
                       dim3(kQJ * kQK), 0, cx.stream, g, x, (int)d[0], (int)d[1], (int)d[2], sq, sqp, p, (int)tiles_k);
    check_launch();
}

void precond_apply(Ctx &cx, const long long d[3], const float *in0, const float *in1, float *out0, float *out1,
                   float2 *work) {
    const int n1 = (int)d[0], n2 = (int)d[1], n3 = (int)d[2];
    const long long n = d[0] * d[1] * d[2];
    const long long plane = (long long)n2 * n3;
    // Two real slices go through one complex transform (the filter is real and even, so
    // the two stay apart): the slices of in0 and in1, or, for one component, slice s
    // and slice s + ceil(n1/2) of in0 (an odd n1 leaves the middle slice alone).
    int planes = n1, paired = 1 << 30;
    const float *i1 = in1;
    float *o1 = out1;
    if (!in1 && n1 > 1) {
        planes = (n1 + 1) / 2;
        paired = n1 - planes;
        i1 = in0 + planes * plane;
        o1 = out0 + planes * plane;
    }
    float2 *a = work;
    if (!a) {
        cx.work[2].ensure(sizeof(float2) * (size_t)n);
        a = cx.work[2].as<float2>();
    }
    // rows (n3), real or real-pair in
    {
        AxisPass p;
        p.in = in0; p.in2 = i1;
        p.ld_mode = i1 ? LD_REAL_PAIR : LD_REAL;
        p.c1_in2 = paired;
        p.is1 = plane; p.is2 = n3; p.isn = 1;
        p.out = a; p.os1 = plane; p.os2 = n3; p.osn = 1;
        p.c1 = planes; p.c2 = n2; p.ci = 1;
        run_axis(cx, get_fft_plan(cx, n3, -1), p);
    }
    {
        AxisPass p;
        p.in = a; p.out = a;
        p.is1 = p.os1 = plane; p.isi = p.osi = 1; p.isn = p.osn = n3;
        p.c1 = planes; p.c2 = 1; p.ci = n3;
        run_axis(cx, get_fft_plan(cx, n2, -1), p);
    }
    hipLaunchKernelGGL(k_ramp_filter, dim3(vgrid((long long)planes * plane)), dim3(256), 0, cx.stream, a,
                       (float2 *)nullptr, planes, n2, n3);
    check_launch();
    {
        AxisPass p;
        p.in = a; p.out = a;
        p.is1 = p.os1 = plane; p.isi = p.osi = 1; p.isn = p.osn = n3;
        p.c1 = planes; p.c2 = 1; p.ci = n3;
        run_axis(cx, get_fft_plan(cx, n2, +1), p);
    }
    {
        AxisPass p;
        p.in = a; p.is1 = plane; p.is2 = n3; p.isn = 1;
        p.out = out0; p.out2 = o1;
        p.st_mode = o1 ? ST_REAL_PAIR : ST_REAL;
        p.c1_out2 = paired;
        p.os1 = plane; p.os2 = n3; p.osn = 1;
        p.c1 = planes; p.c2 = n2; p.ci = 1;
        p.scale = (float)(1.0 / ((double)n2 * (double)n3));
        run_axis(cx, get_fft_plan(cx, n3, +1), p);
    }
    (void)n;
}

void axpby3(Ctx &cx, long long n, float a, CVol3 x, float b, Vol3 y) {
    hipLaunchKernelGGL(k_axpby3, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, a, x, b, y);
    check_launch();
}

void copy3(Ctx &cx, long long n, CVol3 x, Vol3 y) {
    hipLaunchKernelGGL(k_copy3, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, x, y);
    check_launch();
}

void fill(Ctx &cx, long long n, float *y, float v) {
    hipLaunchKernelGGL(k_fill, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, y, v);
    check_launch();
}

void scale_to(Ctx &cx, long long n, const float *x, float s, float *y) {
    hipLaunchKernelGGL(k_scale_to, dim3(vgrid(n)), dim3(256), 0, cx.stream, n, x, s, y);
    check_launch();
}

}  // namespace lam
