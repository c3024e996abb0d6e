// Toeplitz normal operator: closed-form FP64 kernel build, packed real FFTs with
// pruned zero padding, 3x3 spectral multiply, Parseval quadratic forms.
#include <cmath>
#include <cstdlib>

#include "fft.h"
#include "toeplitz.h"

namespace lam {

namespace {

size_t next_smooth(size_t n) {
    for (;; ++n) {
        size_t r = n;
        for (size_t p : {2, 3, 5}) while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

// sin(pi v) / sin(pi v / n); the removable singularity at v/n = m gives n (-1)^(m(n-1))
__device__ __forceinline__ double dirichlet(double v, int n) {
    const double den = sinpi(v / (double)n);
    if (fabs(den) < 1e-10) {
        const long long m = llrint(v / (double)n);
        return ((m * (long long)(n - 1)) & 1LL) ? -(double)n : (double)n;
    }
    return sinpi(v) / den;
}

// Rotations that step the D_R arguments pi dv and pi dv / R of one projection
// by one dy: (cos, sin) of pi vy and of pi vy / R.
struct DyStep {
    double c1, s1, c2, s2;
};

constexpr int kToepDy = 8;   // dy offsets per thread

// Closed-form kernel T_ik(D) = scale sum_p d_i d_k D_C(D.u_p) D_R(D.v_p) on the offsets
// D in [-(n-1), n-1]^3, added at D mod L (6 unique (i, k) pairs in 3 complex slots).
// T is even in D (so are the Dirichlet kernels): only dx >= 0 is computed and each
// value is added at D and -D; on the dx = 0 plane the half dz > 0 or (dz = 0, dy > 0)
// is mirrored and the centre added once, so every offset is written exactly once.
// D_C(D.u) does not depend on dy, and D.v is linear in dy: a thread takes kToepDy
// consecutive dy, evaluates D_C once per projection and steps the sines of D_R by
// rotation (error ~kToepDy ulp).
__global__ __launch_bounds__(256) void k_toep_build(float2 *__restrict__ K, const Geo *__restrict__ geo,
                                                    const DyStep *__restrict__ step, int P, int R, int C, int n1,
                                                    int n2, int n3, int L1, int L2, int L3, double scale) {
    const int e2 = 2 * n2 - 1;
    const int nyc = (e2 + kToepDy - 1) / kToepDy;
    const long long total = (long long)(2 * n1 - 1) * n3 * nyc;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int yc = (int)(id % nyc);
        const int dx = (int)((id / nyc) % n3);
        const int dz = (int)(id / ((long long)nyc * n3)) - (n1 - 1);
        const int dy0 = -(n2 - 1) + yc * kToepDy;
        const int ny = min(kToepDy, (n2 - 1) - dy0 + 1);
        if (dx == 0 && dz < 0) continue;                                  // mirrored from dz > 0
        if (dx == 0 && dz == 0 && dy0 + ny - 1 < 0) continue;            // mirrored from dy > 0
        double acc[kToepDy][6];
#pragma unroll
        for (int k = 0; k < kToepDy; ++k)
#pragma unroll
            for (int q = 0; q < 6; ++q) acc[k][q] = 0.0;
        for (int p = 0; p < P; ++p) {
            const Geo g = geo[p];
            const DyStep st = step[p];
            const double Dc = dirichlet(dx * g.ux + dz * g.uz, C);
            const double dv0 = dx * g.vx + dy0 * g.vy + dz * g.vz;
            double s1, c1, s2, c2;
            sincospi(dv0, &s1, &c1);
            sincospi(dv0 / (double)R, &s2, &c2);
#pragma unroll
            for (int k = 0; k < kToepDy; ++k) {
                if (k < ny) {
                    double Dr;
                    if (fabs(s2) < 1e-10) {
                        const long long m = llrint((dv0 + k * g.vy) / (double)R);
                        Dr = ((m * (long long)(R - 1)) & 1LL) ? -(double)R : (double)R;
                    } else {
                        Dr = s1 / s2;
                    }
                    const double D = Dc * Dr;
                    const double e0 = g.d0 * D, e1v = g.d1 * D, e2v = g.d2 * D;
                    acc[k][0] += g.d0 * e0; acc[k][1] += g.d0 * e1v; acc[k][2] += g.d0 * e2v;
                    acc[k][3] += g.d1 * e1v; acc[k][4] += g.d1 * e2v; acc[k][5] += g.d2 * e2v;
                    const double t1 = c1 * st.c1 - s1 * st.s1, t2 = c2 * st.c2 - s2 * st.s2;
                    s1 = s1 * st.c1 + c1 * st.s1;
                    s2 = s2 * st.c2 + c2 * st.s2;
                    c1 = t1;
                    c2 = t2;
                }
            }
        }
        auto add = [&](int z, int y, int x, const double *a) {
            const int pz = z < 0 ? z + L1 : z, py = y < 0 ? y + L2 : y, px = x < 0 ? x + L3 : x;
            const long long pos = (((long long)pz * L2 + py) * L3 + px) * 3;
            float2 k0 = K[pos], k1 = K[pos + 1], k2 = K[pos + 2];
            k0.x += (float)(a[0] * scale); k0.y += (float)(a[1] * scale);
            k1.x += (float)(a[2] * scale); k1.y += (float)(a[3] * scale);
            k2.x += (float)(a[4] * scale); k2.y += (float)(a[5] * scale);
            K[pos] = k0; K[pos + 1] = k1; K[pos + 2] = k2;
        };
#pragma unroll
        for (int k = 0; k < kToepDy; ++k) {
            if (k >= ny) break;
            const int dy = dy0 + k;
            if (dx == 0 && dz == 0 && dy < 0) continue;
            add(dz, dy, dx, acc[k]);
            if (dx != 0 || dz != 0 || dy != 0) add(-dz, -dy, -dx, acc[k]);
        }
    }
}

// X0, X1 from the packed spectrum Za = FFT(x0 + i x1) at f and -f; X2 = Zb(f).
__device__ __forceinline__ void unpack(float2 a, float2 b, float2 &X0, float2 &X1) {
    X0 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    const float2 t = make_float2(a.x - b.x, a.y + b.y);   // a - conj(b)
    X1 = make_float2(0.5f * t.y, -0.5f * t.x);            // t / (2i)
}

// The spectra are walked by rows (z, y) of length L3: the row and its mirror row
// -f are found with one 32-bit division per row, never per element.
struct SpecRow {
    long long base, mbase;
};

__device__ __forceinline__ SpecRow spec_row(int r, int L1, int L2, int L3) {
    const int z = r / L2, y = r - z * L2;
    const int mz = z ? L1 - z : 0, my = y ? L2 - y : 0;
    return {(long long)r * L3, ((long long)mz * L2 + my) * L3};
}

__global__ __launch_bounds__(256) void k_toep_mul(const float2 *__restrict__ za, const float2 *__restrict__ zb,
                                                  const float2 *__restrict__ K, float2 *__restrict__ wa,
                                                  float2 *__restrict__ wb, int L1, int L2, int L3) {
    const int rows = L1 * L2;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
      const SpecRow sr = spec_row(r, L1, L2, L3);
      for (int x = threadIdx.x; x < L3; x += blockDim.x) {
        const long long id = sr.base + x;
        float2 X0, X1;
        unpack(za[id], za[sr.mbase + (x ? L3 - x : 0)], X0, X1);
        const float2 X2 = zb[id];
        const float2 k0 = K[3 * id], k1 = K[3 * id + 1], k2 = K[3 * id + 2];
        float2 Y0, Y1, Y2;
        Y0.x = k0.x * X0.x + k0.y * X1.x + k1.x * X2.x;
        Y0.y = k0.x * X0.y + k0.y * X1.y + k1.x * X2.y;
        Y1.x = k0.y * X0.x + k1.y * X1.x + k2.x * X2.x;
        Y1.y = k0.y * X0.y + k1.y * X1.y + k2.x * X2.y;
        Y2.x = k1.x * X0.x + k2.x * X1.x + k2.y * X2.x;
        Y2.y = k1.x * X0.y + k2.x * X1.y + k2.y * X2.y;
        wa[id] = make_float2(Y0.x - Y1.y, Y0.y + Y1.x);
        wb[id] = Y2;
      }
    }
}

constexpr int kMaxQuadGrids = 8;
struct QuadSpec {
    const float2 *K[kMaxQuadGrids];
    int G;
};

__global__ __launch_bounds__(256) void k_toep_quad(const float2 *__restrict__ za, const float2 *__restrict__ zb,
                                                   QuadSpec s, int L1, int L2, int L3, double *partials) {
    double acc[kMaxQuadGrids * 3];
#pragma unroll
    for (int q = 0; q < kMaxQuadGrids * 3; ++q) acc[q] = 0.0;
    // The spectra of real volumes are Hermitian (X(-k) = conj X(k)) and K is real and
    // even, so the term at -k equals the term at k: the sum runs over one of each pair
    // of mirrored points with weight 2, and once over the points that are their own
    // mirror (rows with r = mirror(r) at x = 0 and x = L3/2), which halves the reads.
    const int rows = L1 * L2;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
      const SpecRow sr = spec_row(r, L1, L2, L3);
      const long long mr = sr.mbase / L3;
      if (mr < r) continue;  // the mirror row carries these terms
      const bool self = mr == r;
      const int xend = self ? L3 / 2 + 1 : L3;
      for (int x = threadIdx.x; x < xend; x += blockDim.x) {
        const long long id = sr.base + x;
        const double w = (self && (x == 0 || 2 * x == L3)) ? 1.0 : 2.0;
        float2 X0, X1;
        unpack(za[id], za[sr.mbase + (x ? L3 - x : 0)], X0, X1);
        const float2 X2 = zb[id];
#pragma unroll
        for (int g = 0; g < kMaxQuadGrids; ++g) {
            if (g < s.G) {
                const float2 *K = s.K[g];
                const float2 k0 = K[3 * id], k1 = K[3 * id + 1], k2 = K[3 * id + 2];
                // Re(conj(X_i) Y_i) with Y = K X, K real symmetric
                const double x0x = X0.x, x0y = X0.y, x1x = X1.x, x1y = X1.y, x2x = X2.x, x2y = X2.y;
                const double y0x = k0.x * x0x + k0.y * x1x + k1.x * x2x, y0y = k0.x * x0y + k0.y * x1y + k1.x * x2y;
                const double y1x = k0.y * x0x + k1.y * x1x + k2.x * x2x, y1y = k0.y * x0y + k1.y * x1y + k2.x * x2y;
                const double y2x = k1.x * x0x + k2.x * x1x + k2.y * x2x, y2y = k1.x * x0y + k2.x * x1y + k2.y * x2y;
                acc[3 * g + 0] += w * (x0x * y0x + x0y * y0y);
                acc[3 * g + 1] += w * (x1x * y1x + x1y * y1y);
                acc[3 * g + 2] += w * (x2x * y2x + x2y * y2y);
            }
        }
      }
    }
    const int nt = 3 * s.G;
#pragma unroll
    for (int q = 0; q < kMaxQuadGrids * 3; ++q) {
        if (q < nt) {
            const double v = block_sum_256(acc[q]);
            if (threadIdx.x == 0) partials[(size_t)q * gridDim.x + blockIdx.x] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *partials, int nblocks, int nterms, double *out) {
    for (int q = 0; q < nterms; ++q) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[(size_t)q * nblocks + b];
        const double v = block_sum_256(acc);
        if (threadIdx.x == 0) out[q] = v;
    }
}

__global__ __launch_bounds__(256) void k_add_c(float2 *a, const float2 *b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a[i] = cadd(a[i], b[i]);
}

unsigned lgrid(long long n) { return grid_for((size_t)n, 256, 1u << 16); }

// The normal operator runs its x transforms and spectral multiply as one fused
// pass (measured: 5.42 -> 3.43 ms per CG iteration on a (720, 144, 720) grid,
// 0.99 -> 0.78 ms on (360, 90, 360)); LAMINO_TOEP_FUSED=0 at operator creation
// selects the separate passes (three transforms each way and a multiply pass).
bool toep_fused() {
    const char *e = std::getenv("LAMINO_TOEP_FUSED");
    return !(e && e[0] == '0');
}

void forward_spectra(Toep &t, CVol3 x) {
    Ctx &cx = *t.cx;
    const int n1 = (int)t.n[0], n2 = (int)t.n[1], n3 = (int)t.n[2];
    const int L1 = t.L[0], L2 = t.L[1], L3 = t.L[2];
    float2 *Z[2] = {t.za.as<float2>(), t.zb.as<float2>()};
    for (int h = 0; h < 2; ++h) {
        {   // x: lines (z < n1, y < n2), real (pair) in, zero-padded to L3
            AxisPass p;
            p.in = x.c[2 * h];
            p.in2 = h == 0 ? x.c[1] : nullptr;
            p.ld_mode = h == 0 ? LD_REAL_PAIR : LD_REAL;
            p.is1 = (long long)n2 * n3; p.is2 = n3; p.isn = 1;
            p.out = Z[h]; p.os1 = (long long)L2 * L3; p.os2 = L3; p.osn = 1;
            p.c1 = n1; p.c2 = n2; p.ci = 1;
            p.nv = n3;
            run_axis(cx, get_fft_plan(cx, L3, -1), p);
        }
        {   // y: lines (z < n1, x < L3)
            AxisPass p;
            p.in = p.out = Z[h];
            p.is1 = p.os1 = (long long)L2 * L3; p.isi = p.osi = 1; p.isn = p.osn = L3;
            p.c1 = n1; p.c2 = 1; p.ci = L3;
            p.nv = n2;
            run_axis(cx, get_fft_plan(cx, L2, -1), p);
        }
        {   // z: lines (y < L2, x < L3)
            AxisPass p;
            p.in = p.out = Z[h];
            p.is2 = p.os2 = L3; p.isi = p.osi = 1; p.isn = p.osn = (long long)L2 * L3;
            p.c1 = 1; p.c2 = L2; p.ci = L3;
            p.nv = n1;
            run_axis(cx, get_fft_plan(cx, L1, -1), p);
        }
    }
}

void inverse_spectra(Toep &t, Vol3 y) {
    Ctx &cx = *t.cx;
    const int n1 = (int)t.n[0], n2 = (int)t.n[1], n3 = (int)t.n[2];
    const int L1 = t.L[0], L2 = t.L[1], L3 = t.L[2];
    float2 *W[2] = {t.wa.as<float2>(), t.wb.as<float2>()};
    for (int h = 0; h < 2; ++h) {
        {   // z: all lines, keep z < n1
            AxisPass p;
            p.in = p.out = W[h];
            p.is2 = p.os2 = L3; p.isi = p.osi = 1; p.isn = p.osn = (long long)L2 * L3;
            p.c1 = 1; p.c2 = L2; p.ci = L3;
            p.ns = n1;
            run_axis(cx, get_fft_plan(cx, L1, +1), p);
        }
        {   // y: lines (z < n1, x < L3), keep y < n2
            AxisPass p;
            p.in = p.out = W[h];
            p.is1 = p.os1 = (long long)L2 * L3; p.isi = p.osi = 1; p.isn = p.osn = L3;
            p.c1 = n1; p.c2 = 1; p.ci = L3;
            p.ns = n2;
            run_axis(cx, get_fft_plan(cx, L2, +1), p);
        }
        {   // x: lines (z < n1, y < n2), keep x < n3, real (pair) out
            AxisPass p;
            p.in = W[h]; p.is1 = (long long)L2 * L3; p.is2 = L3; p.isn = 1;
            p.out = y.c[2 * h];
            p.out2 = h == 0 ? y.c[1] : nullptr;
            p.st_mode = h == 0 ? ST_REAL_PAIR : ST_REAL;
            p.os1 = (long long)n2 * n3; p.os2 = n3; p.osn = 1;
            p.c1 = n1; p.c2 = n2; p.ci = 1;
            p.ns = n3;
            run_axis(cx, get_fft_plan(cx, L3, +1), p);
        }
    }
}

// The normal operator's transforms with the contiguous axis last (forward) and
// first (inverse), so that the x transforms and the spectral multiply between
// them run as one pass over the grid (toep_spectral_x).  Each volume pair
// h = 0 (x0 + i x1) and h = 1 (x2) goes z then y forward, y then z inverse; the
// passes are pruned to the lines and positions the zero padding leaves nonzero
// (z < n1, y < n2, x < n3 on input) or the crop keeps (on output).
void forward_zy(Toep &t, CVol3 x) {
    Ctx &cx = *t.cx;
    const int n1 = (int)t.n[0], n2 = (int)t.n[1], n3 = (int)t.n[2];
    const int L1 = t.L[0], L2 = t.L[1], L3 = t.L[2];
    float2 *Z[2] = {t.za.as<float2>(), t.zb.as<float2>()};
    for (int h = 0; h < 2; ++h) {
        {   // z: lines (y < n2, x < n3), real (pair) in, zero-padded to L1
            AxisPass p;
            p.in = x.c[2 * h];
            p.in2 = h == 0 ? x.c[1] : nullptr;
            p.ld_mode = h == 0 ? LD_REAL_PAIR : LD_REAL;
            p.is2 = n3; p.isi = 1; p.isn = (long long)n2 * n3;
            p.out = Z[h]; p.os2 = L3; p.osi = 1; p.osn = (long long)L2 * L3;
            p.c1 = 1; p.c2 = n2; p.ci = n3;
            p.nv = n1;
            run_axis(cx, get_fft_plan(cx, L1, -1), p);
        }
        {   // y: lines (z < L1, x < n3), zero-padded to L2
            AxisPass p;
            p.in = p.out = Z[h];
            p.is1 = p.os1 = (long long)L2 * L3; p.isi = p.osi = 1; p.isn = p.osn = L3;
            p.c1 = L1; p.c2 = 1; p.ci = n3;
            p.nv = n2;
            run_axis(cx, get_fft_plan(cx, L2, -1), p);
        }
    }
}

void inverse_yz(Toep &t, Vol3 y) {
    Ctx &cx = *t.cx;
    const int n1 = (int)t.n[0], n2 = (int)t.n[1], n3 = (int)t.n[2];
    const int L1 = t.L[0], L2 = t.L[1], L3 = t.L[2];
    float2 *Z[2] = {t.za.as<float2>(), t.zb.as<float2>()};
    for (int h = 0; h < 2; ++h) {
        {   // y: lines (z < L1, x < n3), keep y < n2
            AxisPass p;
            p.in = p.out = Z[h];
            p.is1 = p.os1 = (long long)L2 * L3; p.isi = p.osi = 1; p.isn = p.osn = L3;
            p.c1 = L1; p.c2 = 1; p.ci = n3;
            p.ns = n2;
            run_axis(cx, get_fft_plan(cx, L2, +1), p);
        }
        {   // z: lines (y < n2, x < n3), keep z < n1, real (pair) out
            AxisPass p;
            p.in = Z[h]; p.is2 = L3; p.isi = 1; p.isn = (long long)L2 * L3;
            p.out = y.c[2 * h];
            p.out2 = h == 0 ? y.c[1] : nullptr;
            p.st_mode = h == 0 ? ST_REAL_PAIR : ST_REAL;
            p.os2 = n3; p.osi = 1; p.osn = (long long)n2 * n3;
            p.c1 = 1; p.c2 = n2; p.ci = n3;
            p.ns = n1;
            run_axis(cx, get_fft_plan(cx, L1, +1), p);
        }
    }
}

}  // namespace

int toep_len(long long n) { return (int)next_smooth(std::max<size_t>(1, 2 * (size_t)n - 1)); }

std::unique_ptr<Toep> toep_create(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], bool per_grid) {
    LAM_REQUIRE(ngrids >= 1, LAM_ERR_ARG, "Toeplitz operator needs at least one grid");
    LAM_REQUIRE(!per_grid || ngrids <= kMaxQuadGrids, LAM_ERR_ARG, "too many grids for per-grid kernels");
    for (int a = 0; a < 3; ++a) LAM_REQUIRE(dims[a] >= 1, LAM_ERR_ARG, "invalid volume dimensions");
    auto t = std::make_unique<Toep>();
    t->cx = &cx;
    t->ngrids = ngrids;
    t->per_grid = per_grid;
    for (int a = 0; a < 3; ++a) {
        t->n[a] = dims[a];
        t->L[a] = toep_len(dims[a]);
    }
    const int L1 = t->L[0], L2 = t->L[1], L3 = t->L[2];
    const long long Lt = (long long)L1 * L2 * L3;
    const size_t kbytes = sizeof(float2) * 3 * (size_t)Lt;
    t->khat.alloc(kbytes);
    LAM_HIP(hipMemsetAsync(t->khat.ptr, 0, kbytes, cx.stream));
    for (int g = 0; g < ngrids; ++g) {
        Grid &gr = *grids[g];
        float2 *dst = t->khat.as<float2>();
        if (per_grid) {
            t->khat_g.emplace_back(kbytes);
            dst = t->khat_g.back().as<float2>();
            LAM_HIP(hipMemsetAsync(dst, 0, kbytes, cx.stream));
        }
        const double scale = 1.0 / ((double)gr.R * (double)gr.C * (double)Lt);
        std::vector<DyStep> steps(gr.P);
        for (long long p = 0; p < gr.P; ++p) {
            const double vy = gr.geo_host[p].vy, pi = 3.14159265358979323846;
            steps[p] = {std::cos(pi * vy), std::sin(pi * vy), std::cos(pi * vy / (double)gr.R),
                        std::sin(pi * vy / (double)gr.R)};
        }
        DevBuf dsteps(sizeof(DyStep) * steps.size());
        upload_now(cx, dsteps.ptr, steps.data(), sizeof(DyStep) * steps.size());
        {
            TimedScope ts(cx, "toep_build");
            const long long nthreads =
                (2 * dims[0] - 1) * dims[2] * ((2 * dims[1] - 1 + kToepDy - 1) / kToepDy);
            hipLaunchKernelGGL(k_toep_build, dim3(lgrid(nthreads)), dim3(256), 0, cx.stream, dst, gr.geo.as<Geo>(),
                               dsteps.as<DyStep>(), (int)gr.P, (int)gr.R, (int)gr.C, (int)dims[0], (int)dims[1],
                               (int)dims[2], L1, L2, L3, scale);
            check_launch();
        }
        if (per_grid) {
            fft3d(cx, dst, L1, L2, L3, 3, -1);
            hipLaunchKernelGGL(k_add_c, dim3(lgrid(3 * Lt)), dim3(256), 0, cx.stream, t->khat.as<float2>(), dst,
                               3 * Lt);
            check_launch();
        }
    }
    if (!per_grid) fft3d(cx, t->khat.as<float2>(), L1, L2, L3, 3, -1);
    t->fused = toep_fused();
    for (DevBuf *b : {&t->za, &t->zb}) b->alloc(sizeof(float2) * (size_t)Lt);
    if (!t->fused)
        for (DevBuf *b : {&t->wa, &t->wb}) b->alloc(sizeof(float2) * (size_t)Lt);
    return t;
}

void toep_apply(Toep &t, CVol3 x, Vol3 y, float mu) {
    Ctx &cx = *t.cx;
    if (t.fused) {
        // z and y transforms, then x transform + 3x3 multiply + inverse x transform in
        // one pass (in place), then y and z inverse transforms
        forward_zy(t, x);
        toep_spectral_x(cx, t.za.as<float2>(), t.zb.as<float2>(), t.khat.as<float2>(), t.za.as<float2>(),
                        t.zb.as<float2>(), t.L[0], t.L[1], t.L[2], (int)t.n[2]);
        inverse_yz(t, y);
    } else {
        forward_spectra(t, x);
        {
            TimedScope ts(cx, "toep_mul");
            const unsigned rows = (unsigned)std::min<long long>((long long)t.L[0] * t.L[1], 1 << 20);
            hipLaunchKernelGGL(k_toep_mul, dim3(rows), dim3(256), 0, cx.stream, t.za.as<float2>(), t.zb.as<float2>(),
                               t.khat.as<float2>(), t.wa.as<float2>(), t.wb.as<float2>(), t.L[0], t.L[1], t.L[2]);
            check_launch();
        }
        inverse_spectra(t, y);
    }
    if (mu != 0.f) add_mu_laplacian3(cx, t.n, x, y, mu);
    ++t.applies;
}

constexpr int kQuadBlocks = 1024;
size_t toep_quad_partials() { return (size_t)kQuadBlocks * kMaxQuadGrids * 3; }

void toep_quad_dev(Toep &t, CVol3 x, double *q_dev, double *partials) {
    Ctx &cx = *t.cx;
    LAM_REQUIRE(t.per_grid, LAM_ERR_STATE, "Toeplitz operator was built without per-grid kernels");
    forward_spectra(t, x);
    QuadSpec s;
    s.G = t.ngrids;
    for (int g = 0; g < kMaxQuadGrids; ++g) s.K[g] = g < t.ngrids ? t.khat_g[g].as<float2>() : nullptr;
    const int nb = kQuadBlocks;
    if (!partials) {
        cx.red.ensure(sizeof(double) * toep_quad_partials());
        partials = cx.red.as<double>();
    }
    hipLaunchKernelGGL(k_toep_quad, dim3(nb), dim3(256), 0, cx.stream, t.za.as<float2>(), t.zb.as<float2>(), s,
                       t.L[0], t.L[1], t.L[2], partials);
    check_launch();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, cx.stream, partials, nb, 3 * t.ngrids, q_dev);
    check_launch();
}

void toep_quad(Toep &t, CVol3 x, double *q_host) {
    Ctx &cx = *t.cx;
    cx.scal.ensure(sizeof(double) * 64);
    toep_quad_dev(t, x, cx.scal.as<double>());
    LAM_HIP(hipMemcpyAsync(q_host, cx.scal.ptr, sizeof(double) * 3 * t.ngrids, hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
}

}  // namespace lam
