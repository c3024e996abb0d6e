// oracle/oracle.cpp -- CPU restatement of the reference's vector-laminography
// reconstruction path (lbl-camera/lamino).  TEST INFRASTRUCTURE ONLY: it is the
// checker for the HIP library (tests/, __graft_entry__.smoke()) and the timed
// CPU baseline in bench.py.  Nothing in the product links or calls it.
//
// Every function is templated on the real type T (float / double) and follows
// the reference call structure so that the FP32 instantiation is the
// "reference CPU path" and the FP64 one (NUFFT tolerance 1e-9) is the oracle.
//
//   geometry        include/polar_grid.h:44-75, include/projection.h:36-41
//   NUFFT           include/nufft.h:37-65, include/finufft_plan.h:93-129
//                   (FINUFFT itself is absent; its published exponential-of-
//                   semicircle method, Barnett, Magland & af Klinteberg,
//                   SIAM J. Sci. Comput. 41 (2019), is restated here)
//   FFT / shifts    include/fft.h:34-199, include/fftutils.h:84-168
//   projectors      src/projection.cpp:41-126
//   normal op       src/gradient.cpp:33-127
//   preconditioner  include/precond.h:33-87
//   stencils        src/finitediff.cpp:7-71, include/bregman.h:33-36
//   qGGMRF          src/qggmrf.cpp:25-95
//   solvers         src/conjgrad.cpp:36-117, src/bregman.cpp:38-134,
//                   src/nagopt.cpp:115-209
//   MBIR drivers    src/mbir_cg.cpp:47-128, src/mbir_sb.cpp:46-114,
//                   src/mbir3.cpp:46-135, include/padding.h:93-232
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class T> using cplx = std::complex<T>;
template <class T> using Vec = std::vector<T>;
template <class T> using Vec3 = std::array<std::vector<T>, 3>;

struct Dims {
    size_t n1 = 0, n2 = 0, n3 = 0;
    size_t size() const { return n1 * n2 * n3; }
    size_t idx(size_t i, size_t j, size_t k) const { return (i * n2 + j) * n3 + k; }
    bool operator==(const Dims &o) const { return n1 == o.n1 && n2 == o.n2 && n3 == o.n3; }
};

static const double kPi = 3.14159265358979323846264338327950288;

// --------------------------------------------------------------------------
// 1-D complex FFT: recursive mixed radix for smooth sizes, Bluestein (power-of-
// two inner length) for sizes with a prime factor above 13.  Unnormalised,
// X[k] = sum_n x[n] exp(sign * 2 pi i n k / N) -- FFTW's FORWARD is sign -1
// and BACKWARD is sign +1 (include/fft.h:57-111).
// --------------------------------------------------------------------------
template <class T> struct Fft1 {
    size_t n = 0;
    int sign = -1;
    std::vector<int> factors;           // radix sequence (smooth path)
    std::vector<cplx<T>> roots;         // exp(sign 2 pi i m / n), m < n
    std::vector<std::vector<cplx<T>>> tw;  // per stage: exp(sign 2 pi i t k / (Ns r)), [t][k]
    bool bluestein = false;
    size_t m = 0;                       // Bluestein inner length (power of two)
    std::unique_ptr<Fft1<T>> inner_f, inner_b;
    std::vector<cplx<T>> chirp, bhat;   // Bluestein tables

    static std::vector<int> factorize(size_t n, bool &smooth) {
        std::vector<int> f;
        smooth = true;
        size_t r = n;
        for (int p : {4, 2, 3, 5, 7, 11, 13}) {
            while (r % p == 0) { f.push_back(p); r /= p; }
        }
        if (r != 1) smooth = false;
        return f;
    }

    Fft1(size_t n_, int sign_) : n(n_), sign(sign_) {
        bool smooth;
        factors = factorize(n, smooth);
        if (n <= 1) { factors.clear(); smooth = true; }
        if (smooth) {
            roots.resize(n);
            for (size_t i = 0; i < n; ++i) {
                long double a = (long double)sign * 2.0L * (long double)kPi * (long double)i / (long double)n;
                roots[i] = cplx<T>((T)std::cos(a), (T)std::sin(a));
            }
            size_t Ns = 1;
            for (int r : factors) {
                std::vector<cplx<T>> t((size_t)r * Ns);
                for (int q = 0; q < r; ++q)
                    for (size_t k = 0; k < Ns; ++k) t[q * Ns + k] = roots[(q * k * (n / (Ns * r))) % n];
                tw.push_back(std::move(t));
                Ns *= r;
            }
            return;
        }
        bluestein = true;
        m = 1;
        while (m < 2 * n - 1) m <<= 1;
        inner_f.reset(new Fft1<T>(m, -1));
        inner_b.reset(new Fft1<T>(m, +1));
        chirp.resize(n);
        for (size_t i = 0; i < n; ++i) {
            // exp(sign * pi i * i^2 / n), with i^2 reduced mod 2n exactly
            unsigned long long q = (unsigned long long)i * i % (2ull * n);
            long double a = (long double)sign * (long double)kPi * (long double)q / (long double)n;
            chirp[i] = cplx<T>((T)std::cos(a), (T)std::sin(a));
        }
        std::vector<cplx<T>> b(m, cplx<T>(0));
        for (size_t i = 0; i < n; ++i) {
            b[i] = std::conj(chirp[i]);
            if (i) b[m - i] = std::conj(chirp[i]);
        }
        bhat.resize(m);
        std::vector<cplx<T>> work(m);
        inner_f->exec(b.data(), bhat.data(), work.data());
        for (auto &v : bhat) v /= (T)m;
    }

    // One radix-r pass of the iterative (self-sorting) transform:
    // y[(j/Ns) Ns r + k + t Ns] = sum_q W_r^{qt} (W_{Ns r}^{qk} x[j + q n/r]), k = j mod Ns.
    void pass(const cplx<T> *x, cplx<T> *y, int r, size_t Ns, const cplx<T> *t) const {
        const size_t step = n / r;
        const T s = (T)sign;
        for (size_t j = 0; j < step; ++j) {
            const size_t k = j % Ns;
            const size_t o = (j / Ns) * Ns * r + k;
            cplx<T> v[16];
            v[0] = x[j];
            for (int q = 1; q < r; ++q) v[q] = x[j + q * step] * t[q * Ns + k];
            if (r == 2) {
                y[o] = v[0] + v[1];
                y[o + Ns] = v[0] - v[1];
            } else if (r == 4) {
                const cplx<T> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3];
                const cplx<T> d = (v[1] - v[3]) * cplx<T>(0, s);
                y[o] = a + c;
                y[o + Ns] = b + d;
                y[o + 2 * Ns] = a - c;
                y[o + 3 * Ns] = b - d;
            } else if (r == 3) {
                const cplx<T> a = v[1] + v[2];
                const cplx<T> b = v[0] - a * (T)0.5;
                const cplx<T> c = (v[1] - v[2]) * cplx<T>(0, s * (T)0.86602540378443864676);
                y[o] = v[0] + a;
                y[o + Ns] = b + c;
                y[o + 2 * Ns] = b - c;
            } else {
                for (int q = 0; q < r; ++q) {
                    cplx<T> acc = v[0];
                    for (int p = 1; p < r; ++p) acc += v[p] * roots[((size_t)(p * q) % r) * (n / r)];
                    y[o + q * Ns] = acc;
                }
            }
        }
    }

    // out-of-place transform of a contiguous length-n vector (work: n scratch, m for Bluestein)
    void exec(const cplx<T> *in, cplx<T> *out, cplx<T> *work) const {
        if (n == 1) { out[0] = in[0]; return; }
        if (!bluestein) {
            const int S = (int)factors.size();
            // ping-pong so that the last pass lands in out
            cplx<T> *bufs[2] = {out, work};
            int cur = (S % 2 == 0) ? 0 : 1;
            std::copy(in, in + n, bufs[cur]);
            size_t Ns = 1;
            for (int st = 0; st < S; ++st) {
                pass(bufs[cur], bufs[cur ^ 1], factors[st], Ns, tw[st].data());
                cur ^= 1;
                Ns *= factors[st];
            }
            return;
        }
        std::vector<cplx<T>> a(m, cplx<T>(0)), A(m), w2(m);
        for (size_t i = 0; i < n; ++i) a[i] = in[i] * chirp[i];
        inner_f->exec(a.data(), A.data(), w2.data());
        for (size_t i = 0; i < m; ++i) A[i] *= bhat[i];
        inner_b->exec(A.data(), a.data(), w2.data());
        for (size_t i = 0; i < n; ++i) out[i] = a[i] * chirp[i];
    }
    void exec(const cplx<T> *in, cplx<T> *out) const {
        std::vector<cplx<T>> w(std::max(n, m) + 1);
        exec(in, out, w.data());
    }
};

// Transform along one axis of a row-major [n1][n2][n3] complex array, in place.
template <class T>
void fft_axis(cplx<T> *data, const Dims &d, int axis, int sign, bool parallel = true) {
    size_t len = axis == 0 ? d.n1 : axis == 1 ? d.n2 : d.n3;
    size_t stride = axis == 0 ? d.n2 * d.n3 : axis == 1 ? d.n3 : 1;
    size_t nlines = d.size() / len;
    Fft1<T> plan(len, sign);
#pragma omp parallel if (parallel)
    {
        std::vector<cplx<T>> a(len), b(len), w(std::max(len, plan.m) + 1);
#pragma omp for schedule(static)
        for (long long line = 0; line < (long long)nlines; ++line) {
            size_t base;
            if (axis == 2) base = (size_t)line * len;
            else if (axis == 1) base = ((size_t)line / d.n3) * d.n2 * d.n3 + ((size_t)line % d.n3);
            else base = (size_t)line;
            for (size_t i = 0; i < len; ++i) a[i] = data[base + i * stride];
            plan.exec(a.data(), b.data(), w.data());
            for (size_t i = 0; i < len; ++i) data[base + i * stride] = b[i];
        }
    }
}

// fft::fft2 / ifft2: batched 2-D over (n2, n3), unnormalised
// (include/fft.h:34-114).  The reference runs FFTW single-threaded.
template <class T> void fft2(cplx<T> *data, const Dims &d, int sign, bool parallel) {
    fft_axis(data, d, 2, sign, parallel);
    fft_axis(data, d, 1, sign, parallel);
}

// fft::fftshift2 moves left by floor(n/2); ifftshift2 by ceil(n/2)
// (include/fftutils.h:84-103, 147-168).
template <class T> void shift2(std::vector<cplx<T>> &a, const Dims &d, bool inverse) {
    size_t d1 = d.n2 / 2 + (inverse && (d.n2 % 2) ? 1 : 0);
    size_t d2 = d.n3 / 2 + (inverse && (d.n3 % 2) ? 1 : 0);
    std::vector<cplx<T>> out(a.size());
#pragma omp parallel for collapse(2)
    for (long long i = 0; i < (long long)d.n1; ++i)
        for (long long j = 0; j < (long long)d.n2; ++j)
            for (size_t k = 0; k < d.n3; ++k)
                out[d.idx(i, j, k)] = a[d.idx(i, (j + d1) % d.n2, (k + d2) % d.n3)];
    a.swap(out);
}

// --------------------------------------------------------------------------
// Geometry.  One plane of R x C non-uniform points per projection p with
// in-plane basis u_p = (cos g, 0, -sin g), v_p = (cos a sin g, -sin a, cos a cos g)
// (include/polar_grid.h:63-72; the reference has g_p = gamma for all p and
// a_p = theta_p), and beam direction d_p = beam_dir_vector(g_p, a_p)
// (include/projection.h:36-41).
// --------------------------------------------------------------------------
template <class T> struct Grid {
    size_t P = 0, R = 0, C = 0;
    std::vector<T> alpha, gamma;   // per projection
    std::vector<T> x, y, z;        // [P][R][C] coordinates (the reference's PolarGrid fields)
    size_t npts() const { return P * R * C; }
    std::array<T, 3> dir(size_t p) const {
        T g = gamma[p], a = alpha[p];
        return {std::sin(g) * std::sin(a), std::cos(a), std::cos(g) * std::sin(a)};
    }
};

template <class T>
Grid<T> make_grid(const double *alpha, const double *gamma, size_t P, size_t R, size_t C) {
    Grid<T> g;
    g.P = P; g.R = R; g.C = C;
    g.alpha.assign(alpha, alpha + P);
    g.gamma.assign(gamma, gamma + P);
    g.x.resize(P * R * C); g.y.resize(P * R * C); g.z.resize(P * R * C);
    T dX = (T)((2 * kPi) / (T)C);   // double quotient rounded to T, as polar_grid.h:61-62
    T dY = (T)((2 * kPi) / (T)R);
#pragma omp parallel for collapse(2)
    for (long long i = 0; i < (long long)P; ++i)
        for (long long j = 0; j < (long long)R; ++j)
            for (size_t k = 0; k < C; ++k) {
                T qX = (T)(k * dX - kPi + dX / 2);
                T qY = (T)(j * dY - kPi + dY / 2);
                T ga = g.gamma[i], th = g.alpha[i];
                size_t id = (i * R + j) * C + k;
                g.z[id] = -qX * std::sin(ga) + qY * std::cos(th) * std::cos(ga);
                g.y[id] = -qY * std::sin(th);
                g.x[id] = qX * std::cos(ga) + qY * std::cos(th) * std::sin(ga);
            }
    return g;
}

// --------------------------------------------------------------------------
// NUFFT, 3-D, modes = volume dims (n1,n2,n3) <-> (z,y,x); mode index
// i - floor(n/2) (FINUFFT modeord 0, include/nufft.h:41-43,56-58).
//   type 2 (iflag -1): c_p = sum_k f_k exp(-i k.x_p)
//   type 1 (iflag +1): f_k = sum_p c_p exp(+i k.x_p)
// --------------------------------------------------------------------------
static size_t next_smooth(size_t n) {
    for (;; ++n) {
        size_t r = n;
        for (size_t p : {2, 3, 5}) while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

template <class T> struct Nufft {
    Dims modes;      // n1 (z), n2 (y), n3 (x)
    size_t nf[3];    // fine grid z, y, x
    int w;
    double beta;
    std::vector<double> corr[3];  // 1/psi_hat(k) per axis, indexed by array index

    Nufft(const Dims &d, double eps) : modes(d) {
        w = (int)std::ceil(-std::log10(eps / 10.0));
        if (w < 2) w = 2;
        if (w > 16) w = 16;
        beta = 2.30 * w;
        size_t nn[3] = {d.n1, d.n2, d.n3};
        for (int a = 0; a < 3; ++a) {
            size_t f = std::max<size_t>(2 * nn[a], 2 * (size_t)w);
            f = next_smooth(f);
            if (f % 2) f = next_smooth(f + 1);
            nf[a] = f;
            corr[a].resize(nn[a]);
            // psi_hat(k) = (w/2) int_{-1}^{1} phi(z) cos(pi k w z / nf) dz, Gauss-Legendre
            const int nq = 200;
            std::vector<double> xq(nq), wq(nq);
            gauss_legendre(nq, xq.data(), wq.data());
            for (size_t i = 0; i < nn[a]; ++i) {
                long long k = (long long)i - (long long)(nn[a] / 2);
                double s = 0;
                for (int q = 0; q < nq; ++q)
                    s += wq[q] * phi(xq[q]) * std::cos(kPi * (double)k * w * xq[q] / (double)f);
                corr[a][i] = 1.0 / (0.5 * w * s);
            }
        }
    }
    double phi(double z) const {
        if (std::fabs(z) >= 1.0) return 0.0;
        return std::exp(beta * (std::sqrt(1.0 - z * z) - 1.0));
    }
    static void gauss_legendre(int n, double *x, double *wt) {
        for (int i = 0; i < n; ++i) {
            double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0;
            for (int it = 0; it < 100; ++it) {
                double p1 = 1, p2 = 0;
                for (int j = 1; j <= n; ++j) {
                    double p3 = p2; p2 = p1;
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
    // kernel weights and first cell for one coordinate on axis a
    void weights(T coord, int a, long long &i0, T *wts) const {
        double t = (double)coord * (double)nf[a] / (2 * kPi);
        double fl = std::floor(t / (double)nf[a]);
        t -= fl * (double)nf[a];
        i0 = (long long)std::ceil(t - 0.5 * w);
        for (int m = 0; m < w; ++m) {
            double z = ((double)(i0 + m) - t) * 2.0 / w;
            wts[m] = (T)phi(z);
        }
    }

    // type 2 on the points of grid g: f -> c
    void type2(const cplx<T> *f, const Grid<T> &g, cplx<T> *c) const {
        size_t nfz = nf[0], nfy = nf[1], nfx = nf[2];
        Dims fd{nfz, nfy, nfx};
        std::vector<cplx<T>> G(fd.size(), cplx<T>(0));
        for (size_t i = 0; i < modes.n1; ++i) {
            long long kz = (long long)i - (long long)(modes.n1 / 2);
            size_t gz = (size_t)((kz + (long long)nfz) % (long long)nfz);
            for (size_t j = 0; j < modes.n2; ++j) {
                long long ky = (long long)j - (long long)(modes.n2 / 2);
                size_t gy = (size_t)((ky + (long long)nfy) % (long long)nfy);
                for (size_t k = 0; k < modes.n3; ++k) {
                    long long kx = (long long)k - (long long)(modes.n3 / 2);
                    size_t gx = (size_t)((kx + (long long)nfx) % (long long)nfx);
                    double cf = corr[0][i] * corr[1][j] * corr[2][k];
                    G[fd.idx(gz, gy, gx)] = f[modes.idx(i, j, k)] * (T)cf;
                }
            }
        }
        fft_axis(G.data(), fd, 2, -1);
        fft_axis(G.data(), fd, 1, -1);
        fft_axis(G.data(), fd, 0, -1);
        size_t M = g.npts();
#pragma omp parallel for schedule(static)
        for (long long p = 0; p < (long long)M; ++p) {
            T wx[16], wy[16], wz[16];
            long long ix, iy, iz;
            weights(g.x[p], 2, ix, wx);
            weights(g.y[p], 1, iy, wy);
            weights(g.z[p], 0, iz, wz);
            cplx<T> acc(0);
            for (int a = 0; a < w; ++a) {
                size_t zz = (size_t)(((iz + a) % (long long)nfz + (long long)nfz) % (long long)nfz);
                for (int b = 0; b < w; ++b) {
                    size_t yy = (size_t)(((iy + b) % (long long)nfy + (long long)nfy) % (long long)nfy);
                    cplx<T> row(0);
                    for (int e = 0; e < w; ++e) {
                        size_t xx = (size_t)(((ix + e) % (long long)nfx + (long long)nfx) % (long long)nfx);
                        row += G[fd.idx(zz, yy, xx)] * wx[e];
                    }
                    acc += row * (wz[a] * wy[b]);
                }
            }
            c[p] = acc;
        }
    }

    // type 1 on the points of grid g: c -> f
    void type1(const cplx<T> *c, const Grid<T> &g, cplx<T> *f) const {
        size_t nfz = nf[0], nfy = nf[1], nfx = nf[2];
        Dims fd{nfz, nfy, nfx};
        std::vector<cplx<T>> G(fd.size(), cplx<T>(0));
        size_t M = g.npts();
        int nth = 1;
#ifdef _OPENMP
        nth = omp_get_max_threads();
#endif
        // per-thread fine grids, reduced afterwards (bounded memory)
        size_t budget = (size_t)8 << 30;
        int ncopy = (int)std::max<size_t>(1, std::min<size_t>(nth, budget / (fd.size() * sizeof(cplx<T>) + 1)));
        std::vector<std::vector<cplx<T>>> priv(ncopy - 1);
        for (auto &v : priv) v.assign(fd.size(), cplx<T>(0));
#pragma omp parallel num_threads(ncopy)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            cplx<T> *dst = tid == 0 ? G.data() : priv[tid - 1].data();
#pragma omp for schedule(static)
            for (long long p = 0; p < (long long)M; ++p) {
                T wx[16], wy[16], wz[16];
                long long ix, iy, iz;
                weights(g.x[p], 2, ix, wx);
                weights(g.y[p], 1, iy, wy);
                weights(g.z[p], 0, iz, wz);
                cplx<T> v = c[p];
                for (int a = 0; a < w; ++a) {
                    size_t zz = (size_t)(((iz + a) % (long long)nfz + (long long)nfz) % (long long)nfz);
                    for (int b = 0; b < w; ++b) {
                        size_t yy = (size_t)(((iy + b) % (long long)nfy + (long long)nfy) % (long long)nfy);
                        cplx<T> vb = v * (wz[a] * wy[b]);
                        for (int e = 0; e < w; ++e) {
                            size_t xx = (size_t)(((ix + e) % (long long)nfx + (long long)nfx) % (long long)nfx);
                            dst[fd.idx(zz, yy, xx)] += vb * wx[e];
                        }
                    }
                }
            }
        }
        for (auto &v : priv) {
#pragma omp parallel for
            for (long long i = 0; i < (long long)G.size(); ++i) G[i] += v[i];
        }
        fft_axis(G.data(), fd, 2, +1);
        fft_axis(G.data(), fd, 1, +1);
        fft_axis(G.data(), fd, 0, +1);
        for (size_t i = 0; i < modes.n1; ++i) {
            long long kz = (long long)i - (long long)(modes.n1 / 2);
            size_t gz = (size_t)((kz + (long long)nfz) % (long long)nfz);
            for (size_t j = 0; j < modes.n2; ++j) {
                long long ky = (long long)j - (long long)(modes.n2 / 2);
                size_t gy = (size_t)((ky + (long long)nfy) % (long long)nfy);
                for (size_t k = 0; k < modes.n3; ++k) {
                    long long kx = (long long)k - (long long)(modes.n3 / 2);
                    size_t gx = (size_t)((kx + (long long)nfx) % (long long)nfx);
                    double cf = corr[0][i] * corr[1][j] * corr[2][k];
                    f[modes.idx(i, j, k)] = G[fd.idx(gz, gy, gx)] * (T)cf;
                }
            }
        }
    }
};

// Brute-force NUDFT in long double (tiny sizes only).
template <class T>
void nudft(int type, const Dims &d, const Grid<T> &g, const cplx<T> *in, cplx<T> *out) {
    size_t M = g.npts();
    if (type == 2) {
        for (size_t p = 0; p < M; ++p) {
            long double re = 0, im = 0;
            for (size_t i = 0; i < d.n1; ++i)
                for (size_t j = 0; j < d.n2; ++j)
                    for (size_t k = 0; k < d.n3; ++k) {
                        long double kz = (long double)i - (long double)(d.n1 / 2);
                        long double ky = (long double)j - (long double)(d.n2 / 2);
                        long double kx = (long double)k - (long double)(d.n3 / 2);
                        long double ph = -(kx * g.x[p] + ky * g.y[p] + kz * g.z[p]);
                        cplx<T> f = in[d.idx(i, j, k)];
                        re += f.real() * std::cos(ph) - f.imag() * std::sin(ph);
                        im += f.real() * std::sin(ph) + f.imag() * std::cos(ph);
                    }
            out[p] = cplx<T>((T)re, (T)im);
        }
    } else {
        for (size_t i = 0; i < d.n1; ++i)
            for (size_t j = 0; j < d.n2; ++j)
                for (size_t k = 0; k < d.n3; ++k) {
                    long double kz = (long double)i - (long double)(d.n1 / 2);
                    long double ky = (long double)j - (long double)(d.n2 / 2);
                    long double kx = (long double)k - (long double)(d.n3 / 2);
                    long double re = 0, im = 0;
                    for (size_t p = 0; p < M; ++p) {
                        long double ph = (kx * g.x[p] + ky * g.y[p] + kz * g.z[p]);
                        cplx<T> c = in[p];
                        re += c.real() * std::cos(ph) - c.imag() * std::sin(ph);
                        im += c.real() * std::sin(ph) + c.imag() * std::cos(ph);
                    }
                    out[d.idx(i, j, k)] = cplx<T>((T)re, (T)im);
                }
    }
}

// --------------------------------------------------------------------------
// Projectors (src/projection.cpp:41-126) and the normal operator
// (src/gradient.cpp:33-127).  `ctx` carries the NUFFT tolerance and whether
// the 2-D detector FFTs run multi-threaded (the reference's FFTW calls are
// single-threaded: include/fft.h has no fftw_plan_with_nthreads).
// --------------------------------------------------------------------------
struct Ctx {
    double eps = 1e-9;
    bool fft2_threads = false;
};

template <class T>
Vec<T> forward(const Ctx &cx, const Vec3<T> &m, const Dims &vd, const Grid<T> &g) {
    Dims pd{g.P, g.R, g.C};
    T scale = (T)(g.R * g.C);
    std::vector<cplx<T>> proj(pd.size(), cplx<T>(0));
    Nufft<T> nu(vd, cx.eps);   // plan per shape (include/finufft_plan_cache.h:50-66)
    for (int i = 0; i < 3; ++i) {
        std::vector<cplx<T>> mc(vd.size()), cc(pd.size());
        for (size_t q = 0; q < vd.size(); ++q) mc[q] = cplx<T>(m[i][q]);
        nu.type2(mc.data(), g, cc.data());
        size_t S = g.R * g.C;
#pragma omp parallel for
        for (long long p = 0; p < (long long)g.P; ++p) {
            T coeff = g.dir(p)[i];
            for (size_t s = 0; s < S; ++s) proj[p * S + s] += cc[p * S + s] * coeff;
        }
    }
    shift2(proj, pd, false);
    fft2(proj.data(), pd, +1, cx.fft2_threads);
    shift2(proj, pd, true);
    Vec<T> out(pd.size());
    for (size_t q = 0; q < out.size(); ++q) out[q] = proj[q].real() / scale;
    return out;
}

template <class T>
Vec3<T> adjoint(const Ctx &cx, const Vec<T> &proj, const Dims &vd, const Grid<T> &g) {
    Dims pd{g.P, g.R, g.C};
    std::vector<cplx<T>> c(pd.size());
    for (size_t q = 0; q < c.size(); ++q) c[q] = cplx<T>(proj[q]);
    shift2(c, pd, false);
    fft2(c.data(), pd, -1, cx.fft2_threads);
    shift2(c, pd, true);
    T scale = (T)(g.npts() / g.P);
    Nufft<T> nu(vd, cx.eps);
    Vec3<T> out;
    size_t S = g.R * g.C;
    for (int i = 0; i < 3; ++i) {
        std::vector<cplx<T>> cc(c.size()), mc(vd.size());
#pragma omp parallel for
        for (long long p = 0; p < (long long)g.P; ++p) {
            T coeff = g.dir(p)[i];
            for (size_t s = 0; s < S; ++s) cc[p * S + s] = c[p * S + s] * coeff;
        }
        nu.type1(cc.data(), g, mc.data());
        out[i].resize(vd.size());
        for (size_t q = 0; q < vd.size(); ++q) out[i][q] = mc[q].real() / scale;
    }
    return out;
}

template <class T>
Vec3<T> sysmat(const Ctx &cx, const Vec3<T> &x, const Dims &vd, const Grid<T> &g) {
    T scale = (T)(g.npts() / g.P);
    Nufft<T> nu(vd, cx.eps);
    size_t M = g.npts(), S = g.R * g.C;
    std::array<std::vector<cplx<T>>, 3> cz;
    for (int i = 0; i < 3; ++i) {
        std::vector<cplx<T>> xc(vd.size());
        for (size_t q = 0; q < vd.size(); ++q) xc[q] = cplx<T>(x[i][q]);
        cz[i].assign(M, cplx<T>(0));
        nu.type2(xc.data(), g, cz[i].data());
    }
    std::array<std::vector<cplx<T>>, 3> rz;
    for (int i = 0; i < 3; ++i) rz[i].assign(M, cplx<T>(0));
#pragma omp parallel for
    for (long long p = 0; p < (long long)g.P; ++p) {
        auto d = g.dir(p);
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                T wgt = d[i] * d[k];
                for (size_t s = 0; s < S; ++s) rz[i][p * S + s] += wgt * cz[k][p * S + s];
            }
    }
    Vec3<T> out;
    for (int i = 0; i < 3; ++i) {
        std::vector<cplx<T>> oc(vd.size());
        nu.type1(rz[i].data(), g, oc.data());
        out[i].resize(vd.size());
        for (size_t q = 0; q < vd.size(); ++q) out[i][q] = oc[q].real() / scale;
    }
    return out;
}

template <class T> double dot(const Vec<T> &a, const Vec<T> &b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s)
    for (long long i = 0; i < (long long)a.size(); ++i) s += (double)a[i] * (double)b[i];
    return s;
}
template <class T> double norm2(const Vec<T> &a) { return std::sqrt(dot(a, a)); }

// --------------------------------------------------------------------------
// Closed-form Toeplitz kernel of the normal operator:
//   T_ik(D) = (1/S) sum_p d_i d_k Dir_C(2pi/C D.u_p) Dir_R(2pi/R D.v_p),
//   Dir_n(t) = sin(n t/2)/sin(t/2), D = (Dx, Dy, Dz) mode offsets.
// sysmat(x)_i[m'] = sum_k sum_m T_ik(m'-m) x_k[m] (exact NUDFT algebra of
// src/gradient.cpp:33-89).  Used only to cross-check.
// --------------------------------------------------------------------------
static double dirichlet(double v, size_t n) {
    // sin(pi v) / sin(pi v / n) with the removable singularity handled
    double den = std::sin(kPi * v / (double)n);
    if (std::fabs(den) < 1e-12) {
        double mm = std::round(v / (double)n);
        double sgn = ((long long)std::llround(mm) * (long long)(n - 1)) % 2 ? -1.0 : 1.0;
        return sgn * (double)n;
    }
    return std::sin(kPi * v) / den;
}

template <class T>
void toeplitz_kernel(const Grid<T> &g, const long long *delta, size_t nd, double *out6) {
    double S = (double)(g.R * g.C);
    for (size_t q = 0; q < nd; ++q) {
        double Dx = (double)delta[3 * q + 0], Dy = (double)delta[3 * q + 1], Dz = (double)delta[3 * q + 2];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (size_t p = 0; p < g.P; ++p) {
            double ga = g.gamma[p], al = g.alpha[p];
            double du = Dx * std::cos(ga) - Dz * std::sin(ga);
            double dv = Dx * std::cos(al) * std::sin(ga) - Dy * std::sin(al) + Dz * std::cos(al) * std::cos(ga);
            double dd = dirichlet(du, g.C) * dirichlet(dv, g.R) / S;
            double d[3] = {std::sin(ga) * std::sin(al), std::cos(al), std::cos(ga) * std::sin(al)};
            acc[0] += d[0] * d[0] * dd; acc[1] += d[0] * d[1] * dd; acc[2] += d[0] * d[2] * dd;
            acc[3] += d[1] * d[1] * dd; acc[4] += d[1] * d[2] * dd; acc[5] += d[2] * d[2] * dd;
        }
        for (int k = 0; k < 6; ++k) out6[6 * q + k] = acc[k];
    }
}

// --------------------------------------------------------------------------
// Ramp preconditioner (include/precond.h:33-87): per n1 slice, 2-D r2c over
// (n2,n3), times sqrt((i/n3)^2 + (y_j/n2)^2), c2r, times 1/(n2 n3).  The
// half-spectrum filter applied through c2r equals the even full-spectrum
// filter on the full complex transform of a real slice.
// --------------------------------------------------------------------------
template <class T> Vec<T> precond(const Vec<T> &in, const Dims &d) {
    std::vector<cplx<T>> a(d.size());
    for (size_t q = 0; q < a.size(); ++q) a[q] = cplx<T>(in[q]);
    fft2(a.data(), d, -1, true);
    int n2 = (int)d.n2, n3 = (int)d.n3;
    std::vector<T> filt(d.n2 * d.n3);
    for (int j = 0; j < n2; ++j) {
        T yf = (j <= n2 / 2) ? (T)j / (T)n2 : (T)(j - n2) / (T)n2;
        for (int i = 0; i < n3; ++i) {
            // the half-spectrum column i' = min(i, n3 - i) carries |x-frequency|
            int ih = i <= n3 / 2 ? i : n3 - i;
            int jh = j;
            T yfh = yf;
            if (i > n3 / 2) {  // Hermitian partner of (j, i) is (-j, n3 - i)
                jh = (n2 - j) % n2;
                yfh = (jh <= n2 / 2) ? (T)jh / (T)n2 : (T)(jh - n2) / (T)n2;
            }
            (void)jh;
            T xf = (T)ih / (T)n3;
            filt[(size_t)j * n3 + i] = std::sqrt(xf * xf + yfh * yfh);
        }
    }
#pragma omp parallel for
    for (long long s = 0; s < (long long)d.n1; ++s)
        for (size_t q = 0; q < d.n2 * d.n3; ++q) a[s * d.n2 * d.n3 + q] *= filt[q];
    fft2(a.data(), d, +1, true);
    double scale = 1.0 / (double)(d.n2 * d.n3);
    Vec<T> out(d.size());
    for (size_t q = 0; q < out.size(); ++q) out[q] = (T)(a[q].real() * (T)scale);
    return out;
}

// --------------------------------------------------------------------------
// Finite differences (src/finitediff.cpp:7-71): interior-only central
// differences / 2; divergence is the negative central divergence.
// --------------------------------------------------------------------------
template <class T> Vec3<T> grad_u(const Vec<T> &u, const Dims &d) {
    Vec3<T> g;
    for (int a = 0; a < 3; ++a) g[a].assign(d.size(), T(0));
    if (d.n1 < 3 || d.n2 < 3 || d.n3 < 3) return g;
#pragma omp parallel for collapse(2)
    for (long long i = 1; i < (long long)d.n1 - 1; ++i)
        for (long long j = 1; j < (long long)d.n2 - 1; ++j)
            for (size_t k = 1; k < d.n3 - 1; ++k) {
                size_t s = d.idx(i, j, k);
                g[0][s] = (T)((u[d.idx(i + 1, j, k)] - u[d.idx(i - 1, j, k)]) / 2.0);
                g[1][s] = (T)((u[d.idx(i, j + 1, k)] - u[d.idx(i, j - 1, k)]) / 2.0);
                g[2][s] = (T)((u[d.idx(i, j, k + 1)] - u[d.idx(i, j, k - 1)]) / 2.0);
            }
    return g;
}

template <class T> Vec<T> divergence(const Vec3<T> &v, const Dims &d) {
    Vec<T> out(d.size(), T(0));
    if (d.n1 < 3 || d.n2 < 3 || d.n3 < 3) return out;
#pragma omp parallel for collapse(2)
    for (long long i = 1; i < (long long)d.n1 - 1; ++i)
        for (long long j = 1; j < (long long)d.n2 - 1; ++j)
            for (size_t k = 1; k < d.n3 - 1; ++k) {
                out[d.idx(i, j, k)] = (T)(-(v[0][d.idx(i + 1, j, k)] - v[0][d.idx(i - 1, j, k)]) / 2.0 -
                                          (v[1][d.idx(i, j + 1, k)] - v[1][d.idx(i, j - 1, k)]) / 2.0 -
                                          (v[2][d.idx(i, j, k + 1)] - v[2][d.idx(i, j, k - 1)]) / 2.0);
            }
    return out;
}

template <class T> Vec<T> laplacian(const Vec<T> &u, const Dims &d) { return divergence(grad_u(u, d), d); }

// --------------------------------------------------------------------------
// qGGMRF prior gradient (src/qggmrf.cpp:25-95), including the dz=-1 face
// weight 0.0532, the delta = 0 branch q/(sigma^q C) and interior-only update.
// --------------------------------------------------------------------------
static const float kQW[27] = {0.0302f, 0.037f, 0.0302f, 0.037f, 0.0523f, 0.037f, 0.0302f, 0.037f, 0.0302f,
                              0.037f,  0.0523f, 0.037f, 0.0532f, 0.f,    0.0523f, 0.037f, 0.0523f, 0.037f,
                              0.0302f, 0.037f, 0.0302f, 0.037f, 0.0523f, 0.037f, 0.0302f, 0.037f, 0.0302f};

template <class T> T dpot(T delta, T sigma_q, T sigma_qp, T p) {
    const double Q = 2.0, C = 0.001;
    T ad = std::abs(delta);
    T t1 = (T)(std::pow((double)ad, Q - (double)p) / (double)sigma_qp);
    T t3 = (T)C + t1;
    if (ad > 0) return (T)((double)(delta / (t3 * sigma_q)) * (Q - ((Q - (double)p) * (double)t1) / (double)t3));
    return (T)(Q / ((double)sigma_q * C));
}

template <class T> void qggmrf(Vec<T> &g, const Vec<T> &x, const Dims &d, T sigma, T p) {
    if (d.n1 < 3 || d.n2 < 3 || d.n3 < 3) return;
    const T sq = (T)std::pow((double)sigma, 2.0);
    const T sqp = (T)std::pow((double)sigma, 2.0 - (double)p);
#pragma omp parallel for collapse(2)
    for (long long i = 1; i < (long long)d.n1 - 1; ++i)
        for (long long j = 1; j < (long long)d.n2 - 1; ++j)
            for (size_t k = 1; k < d.n3 - 1; ++k) {
                T u = x[d.idx(i, j, k)], sum = 0;
                for (int a = -1; a < 2; ++a)
                    for (int b = -1; b < 2; ++b)
                        for (int c = -1; c < 2; ++c) {
                            if (!a && !b && !c) continue;
                            T v = x[d.idx(i + a, j + b, k + c)];
                            T w = (T)kQW[(a + 1) * 9 + (b + 1) * 3 + (c + 1)];
                            sum += w * dpot<T>(v - u, sq, sqp, p);
                        }
                g[d.idx(i, j, k)] += sum;
            }
}

// --------------------------------------------------------------------------
// Solvers
// --------------------------------------------------------------------------
template <class T> using Op = std::function<Vec3<T>(const Vec3<T> &)>;

// src/conjgrad.cpp:36-117 (lambda unused; resetPads is a no-op copy)
template <class T>
Vec3<T> cgsolver(const Op<T> &A, const Vec3<T> &y, const Vec3<T> &x0, const Dims &d, size_t max_iter, T tol,
                 int *iters_done = nullptr) {
    Vec3<T> x = x0, r, z, p;
    Vec3<T> tmp = A(x);
    for (int i = 0; i < 3; ++i) {
        r[i].resize(d.size());
        for (size_t q = 0; q < d.size(); ++q) r[i][q] = y[i][q] - tmp[i][q];
    }
    double rs_old = 0;
    for (int i = 0; i < 3; ++i) {
        z[i] = precond(r[i], d);
        p[i] = z[i];
        rs_old += dot(z[i], r[i]);
    }
    size_t it = 0;
    for (; it < max_iter; ++it) {
        Vec3<T> Ap = A(p);
        double pAp = 0;
        for (int i = 0; i < 3; ++i) pAp += dot(p[i], Ap[i]);
        if (std::fabs(pAp) < 1e-10) break;
        double alpha = rs_old / pAp;
        for (int i = 0; i < 3; ++i)
            for (size_t q = 0; q < d.size(); ++q) {
                x[i][q] += (T)(p[i][q] * (T)alpha);
                r[i][q] -= (T)(Ap[i][q] * (T)alpha);
            }
        double rs_new = 0;
        for (int i = 0; i < 3; ++i) {
            z[i] = precond(r[i], d);
            rs_new += dot(z[i], r[i]);
        }
        if (std::sqrt(rs_new) < (double)tol) { ++it; break; }
        double beta = rs_new / rs_old;
        for (int i = 0; i < 3; ++i)
            for (size_t q = 0; q < d.size(); ++q) p[i][q] = z[i][q] + p[i][q] * (T)beta;
        rs_old = rs_new;
    }
    if (iters_done) *iters_done = (int)it;
    return x;
}

// src/bregman.cpp:38-134
template <class T>
Vec3<T> split_bregman(const Op<T> &A, const Vec3<T> &yT, const Vec3<T> &x0, const Dims &d, T lambda, T mu,
                      size_t outer_max, size_t inner_max, T tol, T xtol) {
    const double EPS = 1e-8;
    Vec3<T> x = x0, x_old = x0;
    std::array<Vec3<T>, 3> dd, bb;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { dd[i][j].assign(d.size(), 0); bb[i][j].assign(d.size(), 0); }
    Op<T> Ap = [&](const Vec3<T> &u) {
        Vec3<T> du = A(u);
        for (int i = 0; i < 3; ++i) {
            Vec<T> l = laplacian(u[i], d);
            for (size_t q = 0; q < d.size(); ++q) du[i][q] += l[q] * mu;
        }
        return du;
    };
    for (size_t it = 0; it < outer_max; ++it) {
        Vec3<T> rhs;
        for (int i = 0; i < 3; ++i) {
            Vec3<T> diff;
            for (int j = 0; j < 3; ++j) {
                diff[j].resize(d.size());
                for (size_t q = 0; q < d.size(); ++q) diff[j][q] = dd[i][j][q] - bb[i][j][q];
            }
            Vec<T> dv = divergence(diff, d);
            rhs[i].resize(d.size());
            for (size_t q = 0; q < d.size(); ++q) rhs[i][q] = yT[i][q] + dv[q] * mu;
        }
        x = cgsolver<T>(Ap, rhs, x_old, d, inner_max, tol);
        std::array<Vec3<T>, 3> gx;
        for (int i = 0; i < 3; ++i) gx[i] = grad_u(x[i], d);
        Vec<T> sk(d.size(), (T)1 * (T)EPS);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (size_t q = 0; q < d.size(); ++q) {
                    T v = gx[i][j][q] + bb[i][j][q];
                    sk[q] += v * v;
                }
        for (auto &v : sk) v = std::sqrt(v);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (size_t q = 0; q < d.size(); ++q) {
                    T tmp = (gx[i][j][q] + bb[i][j][q]) / sk[q];
                    dd[i][j][q] = std::max(sk[q] - lambda / mu, T(0)) * tmp;
                }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (size_t q = 0; q < d.size(); ++q) bb[i][j][q] += gx[i][j][q] - dd[i][j][q];
        double nd = 0;
        for (int i = 0; i < 3; ++i) {
            Vec<T> df(d.size());
            for (size_t q = 0; q < d.size(); ++q) df[q] = x[i][q] - x_old[i][q];
            nd += norm2(df) / (norm2(x_old[i]) + EPS);
        }
        x_old = x;
        if (nd < (double)xtol) break;
    }
    return x;
}

// src/nagopt.cpp:115-209 (vector overload)
template <class T>
Vec3<T> nagopt(const Op<T> &grad, const std::function<double(const Vec3<T> &)> &loss, Vec3<T> &x, const Dims &d,
               size_t max_iters, T lipschitz, T tol, T xtol, size_t max_inner) {
    Vec3<T> xold = x;
    double t = 1, tnew = 1;
    double step = 0.1 / (double)lipschitz, step_prev = step;
    for (size_t iter = 0; iter < max_iters; ++iter) {
        for (size_t inner = 0; inner < max_inner; ++inner) {
            double beta = tnew * (1 / t - 1);
            tnew = 0.5 * (std::sqrt(std::pow(t, 4) + 4 * std::pow(t, 2)) - std::pow(t, 2));
            Vec3<T> y;
            for (int i = 0; i < 3; ++i) {
                y[i].resize(d.size());
                for (size_t q = 0; q < d.size(); ++q) y[i][q] = x[i][q] + (x[i][q] - xold[i][q]) * (T)beta;
            }
            Vec3<T> g = grad(y);
            for (int i = 0; i < 3; ++i)
                for (size_t q = 0; q < d.size(); ++q) x[i][q] = y[i][q] - g[i][q] * (T)step;
            double ex = loss(x), ey = loss(y);
            double gy = 0;
            for (int i = 0; i < 3; ++i) gy += norm2(g[i]);
            gy *= 0.5 * step;
            if (ex > ey + gy) step *= 0.9;
            else {
                step = step_prev;
                t = tnew;
                double es = 0, xs = 0;
                for (int i = 0; i < 3; ++i) {
                    Vec<T> df(d.size());
                    for (size_t q = 0; q < d.size(); ++q) df[q] = x[i][q] - xold[i][q];
                    es += norm2(df);
                    xs += norm2(x[i]);
                }
                double xerr = es / xs;
                if (xerr < (double)xtol) return x;
                xold = x;
                break;
            }
        }
        double e = loss(x);
        if (e < (double)tol) break;
    }
    return x;
}

// src/lipschitz.cpp:30-53: power iteration x <- Ax/|Ax|, L = |Ax|, stop when
// |L_new - L| < tol (the previous estimate is returned then), 0 when |Ax|
// vanishes.  The reference's domain is one Array; here it is the 3-component
// vector the vector solvers iterate on (|.| over all three), with the start
// vector supplied by the caller instead of Array::random_like.
template <class T>
T lipschitz(const Op<T> &A, const Vec3<T> &x0, const Dims &d, size_t max_iter, T tol) {
    auto norm3 = [&](const Vec3<T> &v) {
        double s = 0;
        for (int i = 0; i < 3; ++i) s += dot(v[i], v[i]);
        return (T)std::sqrt(s);
    };
    Vec3<T> x = x0;
    T nx = norm3(x);
    for (int i = 0; i < 3; ++i)
        for (size_t q = 0; q < d.size(); ++q) x[i][q] /= nx;
    T L = 0;
    for (size_t it = 0; it < max_iter; ++it) {
        Vec3<T> Ax = A(x);
        T na = norm3(Ax);
        if (std::abs(na) < (T)1e-12) { L = 0; break; }
        for (int i = 0; i < 3; ++i)
            for (size_t q = 0; q < d.size(); ++q) x[i][q] = Ax[i][q] / na;
        if (std::abs(na - L) < tol) break;
        L = na;
    }
    return L;
}

// --------------------------------------------------------------------------
// MBIR drivers (src/mbir_cg.cpp, src/mbir_sb.cpp, src/mbir3.cpp)
// --------------------------------------------------------------------------
struct Dataset {
    size_t P, R, C;            // projection stack (P, R, C) before padding
    const double *proj;        // [P][R][C]
    const double *alpha;       // per projection
    const double *gamma;       // per projection
};

struct Params {
    int method = 1;            // 1 CG, 2 split-Bregman, 3 NAG+qGGMRF
    size_t max_iters = 50, inner_iters = 3;
    double sigma = 1000, p = 1.2, lambda = 0.1, mu = 10, tol = 1e-5, xtol = 1e-5;
    double pad_factor = 1.4142;
};

// pad rule n + 2*floor(n*f/2) with the float arithmetic of include/padding.h:98
template <class T> static size_t pad_n(size_t n, T f) { return n + 2 * ((size_t)(n * f) / 2); }

// The job set-up every driver shares (src/mbir_cg.cpp:53-97, src/mbir_sb.cpp:51-92,
// src/mbir3.cpp:51-92): padded volume dims, one grid per dataset (projections
// normalised by their max and pad2d-ed), yT = sum of the datasets' adjoints.
template <class T> struct Job {
    Dims od;
    std::vector<Grid<T>> grids;
    Vec3<T> yT;
};

template <class T>
Job<T> mbir_setup(const Ctx &cx, const std::vector<Dataset> &ds, const Dims &rd, const Params &pr) {
    Job<T> job;
    T padding = pr.method == 1 ? (T)((float)pr.pad_factor - 1.0) : (T)(float)pr.pad_factor;
    T vpad = pr.method == 1 ? padding : (T)(padding - 1);
    Dims od{pad_n(rd.n1, vpad), pad_n(rd.n2, vpad), pad_n(rd.n3, vpad)};
    std::vector<Grid<T>> grids;
    Vec3<T> yT;
    for (int i = 0; i < 3; ++i) yT[i].assign(od.size(), 0);
    for (const auto &D : ds) {
        Dims pdim{D.P, D.R, D.C};
        Vec<T> y(pdim.size());
        T mx = -std::numeric_limits<T>::infinity();
        for (size_t q = 0; q < y.size(); ++q) { y[q] = (T)D.proj[q]; mx = std::max(mx, y[q]); }
        for (auto &v : y) v /= mx;
        // pad2d symmetric (include/padding.h:93-131)
        size_t R2 = pad_n(D.R, padding), C2 = pad_n(D.C, padding);
        size_t d1 = (R2 - D.R) / 2, d2 = (C2 - D.C) / 2;
        Vec<T> yp(D.P * R2 * C2, 0);
        for (size_t p = 0; p < D.P; ++p)
            for (size_t r = 0; r < D.R; ++r)
                for (size_t c = 0; c < D.C; ++c) yp[(p * R2 + r + d1) * C2 + c + d2] = y[(p * D.R + r) * D.C + c];
        grids.push_back(make_grid<T>(D.alpha, D.gamma, D.P, R2, C2));
        Vec3<T> a = adjoint(cx, yp, od, grids.back());
        for (int i = 0; i < 3; ++i)
            for (size_t q = 0; q < od.size(); ++q) yT[i][q] += a[i][q];
    }
    job.od = od;
    job.grids = std::move(grids);
    job.yT = std::move(yT);
    return job;
}

// src/mbir3.cpp:86-117: the gradient closure (per dataset sysmat(x) - yT, so
// the summed yT is subtracted G times, then the qGGMRF prior) and the loss
// closure (per dataset and component sqrt(x.Ax - 2 x.yT + yTy)).
template <class T> Vec3<T> mbir3_grad(const Ctx &cx, const Job<T> &job, const Vec3<T> &x, const Params &pr) {
    const Dims &od = job.od;
    Vec3<T> g;
    for (size_t gi = 0; gi < job.grids.size(); ++gi) {
        Vec3<T> t = sysmat(cx, x, od, job.grids[gi]);
        for (int i = 0; i < 3; ++i) {
            if (gi == 0) g[i].assign(od.size(), 0);
            for (size_t q = 0; q < od.size(); ++q) g[i][q] += t[i][q] - job.yT[i][q];
        }
    }
    for (int i = 0; i < 3; ++i) qggmrf<T>(g[i], x[i], od, (T)pr.sigma, (T)pr.p);
    return g;
}

template <class T> double mbir3_loss(const Ctx &cx, const Job<T> &job, const Vec3<T> &x) {
    double yTy = 0;
    for (int i = 0; i < 3; ++i) yTy += dot(job.yT[i], job.yT[i]);
    double e = 0;
    for (size_t gi = 0; gi < job.grids.size(); ++gi) {
        Vec3<T> t = sysmat(cx, x, job.od, job.grids[gi]);
        for (int i = 0; i < 3; ++i) e += std::sqrt(dot(x[i], t[i]) - 2.0 * dot(x[i], job.yT[i]) + yTy);
    }
    return e;
}

template <class T>
Vec3<T> mbir(const Ctx &cx, const std::vector<Dataset> &ds, const Dims &rd, const Params &pr) {
    Job<T> job = mbir_setup<T>(cx, ds, rd, pr);
    const Dims od = job.od;
    const auto &grids = job.grids;
    const Vec3<T> &yT = job.yT;
    Op<T> A = [&](const Vec3<T> &x) {
        Vec3<T> Ax = sysmat(cx, x, od, grids[0]);
        for (size_t g = 1; g < grids.size(); ++g) {
            Vec3<T> t = sysmat(cx, x, od, grids[g]);
            for (int i = 0; i < 3; ++i)
                for (size_t q = 0; q < od.size(); ++q) Ax[i][q] += t[i][q];
        }
        return Ax;
    };
    Vec3<T> x0;
    for (int i = 0; i < 3; ++i) x0[i].assign(od.size(), 0);
    Vec3<T> rec;
    if (pr.method == 1) {
        rec = cgsolver<T>(A, yT, x0, od, pr.max_iters, (T)pr.tol);
    } else if (pr.method == 2) {
        rec = split_bregman<T>(A, yT, x0, od, (T)pr.lambda, (T)pr.mu, pr.max_iters, pr.inner_iters, (T)pr.tol,
                               (T)pr.xtol);
    } else {
        Op<T> grad = [&](const Vec3<T> &x) { return mbir3_grad<T>(cx, job, x, pr); };
        auto loss = [&](const Vec3<T> &x) { return mbir3_loss<T>(cx, job, x); };
        rec = nagopt<T>(grad, loss, x0, od, pr.max_iters, (T)pr.mu, (T)pr.tol, (T)pr.xtol, pr.inner_iters);
    }
    // crop3d symmetric (include/padding.h:206-232)
    Dims cs{od.n1 - rd.n1, od.n2 - rd.n2, od.n3 - rd.n3};
    Vec3<T> out;
    for (int i = 0; i < 3; ++i) {
        out[i].resize(rd.size());
        for (size_t a = 0; a < rd.n1; ++a)
            for (size_t b = 0; b < rd.n2; ++b)
                for (size_t c = 0; c < rd.n3; ++c)
                    out[i][rd.idx(a, b, c)] = rec[i][od.idx(a + cs.n1 / 2, b + cs.n2 / 2, c + cs.n3 / 2)];
    }
    return out;
}

}  // namespace orc

// ==========================================================================
// extern "C" surface for the Python test harness (ctypes).  All arrays are
// double (complex = interleaved re,im); `prec` selects the arithmetic
// (32 -> float, 64 -> double).
// ==========================================================================
using namespace orc;

namespace {
std::string g_err;
template <class F> int guard(F &&f) {
    try { f(); return 0; } catch (const std::exception &e) { g_err = e.what(); return 1; }
}
template <class T> Vec<T> tovec(const double *a, size_t n) { return Vec<T>(a, a + n); }
template <class T> void fromvec(const Vec<T> &v, double *out) { for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i]; }
template <class T> std::vector<cplx<T>> tocvec(const double *a, size_t n) {
    std::vector<cplx<T>> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = cplx<T>((T)a[2 * i], (T)a[2 * i + 1]);
    return v;
}
template <class T> void fromcvec(const std::vector<cplx<T>> &v, double *out) {
    for (size_t i = 0; i < v.size(); ++i) { out[2 * i] = v[i].real(); out[2 * i + 1] = v[i].imag(); }
}
Dims mkd(const long long *d) { return Dims{(size_t)d[0], (size_t)d[1], (size_t)d[2]}; }
}  // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// 1-D FFT over `batch` contiguous vectors of length n (complex, in place)
int orc_fft1d(double *data, long long n, long long batch, int sign) {
    return guard([&] {
        Fft1<double> plan((size_t)n, sign);
        std::vector<cplx<double>> a(n), b(n);
        for (long long q = 0; q < batch; ++q) {
            for (long long i = 0; i < n; ++i) a[i] = cplx<double>(data[2 * (q * n + i)], data[2 * (q * n + i) + 1]);
            plan.exec(a.data(), b.data());
            for (long long i = 0; i < n; ++i) { data[2 * (q * n + i)] = b[i].real(); data[2 * (q * n + i) + 1] = b[i].imag(); }
        }
    });
}

// fft::fftshift2 (inverse=0) / ifftshift2 (inverse=1) on complex [n1][n2][n3]
int orc_shift2(double *data, const long long *dims, int inverse) {
    return guard([&] {
        Dims d = mkd(dims);
        auto v = tocvec<double>(data, d.size());
        shift2(v, d, inverse != 0);
        fromcvec(v, data);
    });
}

int orc_grid_coords(int prec, const double *alpha, const double *gamma, long long P, long long R, long long C,
                    double *x, double *y, double *z) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Grid<T> g = make_grid<T>(alpha, gamma, P, R, C);
            fromvec(g.x, x); fromvec(g.y, y); fromvec(g.z, z);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// NUFFT (type 1 or 2) on the polar points; brute=1 selects the direct NUDFT.
int orc_nufft(int prec, int type, int brute, double eps, const long long *dims, const double *alpha,
              const double *gamma, long long P, long long R, long long C, const double *in, double *out) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Grid<T> g = make_grid<T>(alpha, gamma, P, R, C);
            size_t nin = type == 2 ? d.size() : g.npts(), nout = type == 2 ? g.npts() : d.size();
            auto a = tocvec<T>(in, nin);
            std::vector<cplx<T>> b(nout);
            if (brute) nudft<T>(type, d, g, a.data(), b.data());
            else {
                Nufft<T> nu(d, eps);
                if (type == 2) nu.type2(a.data(), g, b.data()); else nu.type1(a.data(), g, b.data());
            }
            fromcvec(b, out);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_nufft_info(double eps, const long long *dims, long long *nf_out, int *w_out) {
    return guard([&] {
        Nufft<double> nu(mkd(dims), eps);
        for (int a = 0; a < 3; ++a) nf_out[a] = (long long)nu.nf[a];
        *w_out = nu.w;
    });
}

// forward (src/projection.cpp:41-77): m = 3 volumes [3][n1][n2][n3] -> proj [P][R][C]
int orc_forward(int prec, double eps, int fft_threads, const long long *dims, const double *m, const double *alpha,
                const double *gamma, long long P, long long R, long long C, double *proj) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Ctx cx; cx.eps = eps; cx.fft2_threads = fft_threads != 0;
            Grid<T> g = make_grid<T>(alpha, gamma, P, R, C);
            Vec3<T> mv;
            for (int i = 0; i < 3; ++i) mv[i] = tovec<T>(m + i * d.size(), d.size());
            fromvec(forward(cx, mv, d, g), proj);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// adjoint (src/projection.cpp:89-126): proj [P][R][C] -> 3 volumes
int orc_adjoint(int prec, double eps, int fft_threads, const long long *dims, const double *proj,
                const double *alpha, const double *gamma, long long P, long long R, long long C, double *m) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Ctx cx; cx.eps = eps; cx.fft2_threads = fft_threads != 0;
            Grid<T> g = make_grid<T>(alpha, gamma, P, R, C);
            Vec3<T> out = adjoint(cx, tovec<T>(proj, g.npts()), d, g);
            for (int i = 0; i < 3; ++i) fromvec(out[i], m + i * d.size());
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// sysmat summed over G datasets (src/gradient.cpp:33-89; the MBIR A() closure)
int orc_sysmat(int prec, double eps, const long long *dims, const double *x, int G, const long long *P,
               const long long *R, const long long *C, const double *alpha, const double *gamma, double *y) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Ctx cx; cx.eps = eps;
            Vec3<T> xv, acc;
            for (int i = 0; i < 3; ++i) { xv[i] = tovec<T>(x + i * d.size(), d.size()); acc[i].assign(d.size(), 0); }
            size_t off = 0;
            for (int gi = 0; gi < G; ++gi) {
                Grid<T> g = make_grid<T>(alpha + off, gamma + off, P[gi], R[gi], C[gi]);
                off += P[gi];
                Vec3<T> t = sysmat(cx, xv, d, g);
                for (int i = 0; i < 3; ++i)
                    for (size_t q = 0; q < d.size(); ++q) acc[i][q] += t[i][q];
            }
            for (int i = 0; i < 3; ++i) fromvec(acc[i], y + i * d.size());
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_toeplitz_kernel(const double *alpha, const double *gamma, long long P, long long R, long long C,
                        const long long *delta, long long nd, double *out6) {
    return guard([&] {
        Grid<double> g;
        g.P = P; g.R = R; g.C = C;
        g.alpha.assign(alpha, alpha + P);
        g.gamma.assign(gamma, gamma + P);
        toeplitz_kernel(g, delta, (size_t)nd, out6);
    });
}

int orc_precond(int prec, const long long *dims, const double *in, double *out) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            fromvec(precond(tovec<T>(in, d.size()), d), out);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_grad_u(int prec, const long long *dims, const double *u, double *g3) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec3<T> g = grad_u(tovec<T>(u, d.size()), d);
            for (int i = 0; i < 3; ++i) fromvec(g[i], g3 + i * d.size());
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_divergence(int prec, const long long *dims, const double *v3, double *out) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec3<T> v;
            for (int i = 0; i < 3; ++i) v[i] = tovec<T>(v3 + i * d.size(), d.size());
            fromvec(divergence(v, d), out);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_qggmrf(int prec, const long long *dims, double *g, const double *x, double sigma, double p) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec<T> gv = tovec<T>(g, d.size());
            qggmrf<T>(gv, tovec<T>(x, d.size()), d, (T)sigma, (T)p);
            fromvec(gv, g);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// cgsolver on the explicit diagonal operator A = diag(a) (tests/test_cg.cpp:16-54 pattern)
int orc_cg_diag(int prec, const long long *dims, const double *diag3, const double *b3, const double *x03,
                long long max_iter, double tol, double *x3, int *iters) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec3<T> dg, b, x0;
            for (int i = 0; i < 3; ++i) {
                dg[i] = tovec<T>(diag3 + i * d.size(), d.size());
                b[i] = tovec<T>(b3 + i * d.size(), d.size());
                x0[i] = tovec<T>(x03 + i * d.size(), d.size());
            }
            Op<T> A = [&](const Vec3<T> &x) {
                Vec3<T> y;
                for (int i = 0; i < 3; ++i) {
                    y[i].resize(d.size());
                    for (size_t q = 0; q < d.size(); ++q) y[i][q] = dg[i][q] * x[i][q];
                }
                return y;
            };
            Vec3<T> x = cgsolver<T>(A, b, x0, d, (size_t)max_iter, (T)tol, iters);
            for (int i = 0; i < 3; ++i) fromvec(x[i], x3 + i * d.size());
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// Solvers on caller-supplied operators (ctypes callbacks on double arrays
// [3][n]); the oracle's arithmetic runs in T and converts at the callback.
typedef void (*orc_op_cb)(void *user, const double *x3, double *y3);
typedef double (*orc_loss_cb)(void *user, const double *x3);
}  // extern "C"

namespace {
template <class T> Vec3<T> load3(const double *a, size_t n) {
    Vec3<T> v;
    for (int i = 0; i < 3; ++i) v[i] = tovec<T>(a + i * n, n);
    return v;
}
template <class T> void store3(const Vec3<T> &v, double *a) {
    for (int i = 0; i < 3; ++i) fromvec(v[i], a + i * v[i].size());
}
template <class T> Op<T> cb_op(orc_op_cb cb, void *user, size_t n) {
    return [=](const Vec3<T> &x) {
        std::vector<double> in(3 * n), out(3 * n);
        store3(x, in.data());
        cb(user, in.data(), out.data());
        return load3<T>(out.data(), n);
    };
}
std::vector<Dataset> mkds(int G, const long long *P, const long long *R, const long long *C, const double *proj,
                          const double *alpha, const double *gamma) {
    std::vector<Dataset> ds;
    size_t po = 0, ao = 0;
    for (int g = 0; g < G; ++g) {
        ds.push_back(Dataset{(size_t)P[g], (size_t)R[g], (size_t)C[g], proj + po, alpha + ao, gamma + ao});
        po += (size_t)(P[g] * R[g] * C[g]);
        ao += (size_t)P[g];
    }
    return ds;
}
}  // namespace

extern "C" {

int orc_split_bregman_cb(int prec, const long long *dims, orc_op_cb A, void *user, const double *yT3,
                         const double *x03, double lambda, double mu, long long outer_max, long long inner_max,
                         double tol, double xtol, double *out3) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec3<T> x = split_bregman<T>(cb_op<T>(A, user, d.size()), load3<T>(yT3, d.size()),
                                         load3<T>(x03, d.size()), d, (T)lambda, (T)mu, (size_t)outer_max,
                                         (size_t)inner_max, (T)tol, (T)xtol);
            store3(x, out3);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_cgsolver_cb(int prec, const long long *dims, orc_op_cb A, void *user, const double *b3, const double *x03,
                    long long max_iter, double tol, double *out3, int *iters) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            Vec3<T> x = cgsolver<T>(cb_op<T>(A, user, d.size()), load3<T>(b3, d.size()), load3<T>(x03, d.size()),
                                    d, (size_t)max_iter, (T)tol, iters);
            store3(x, out3);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// x3 is the caller's x, updated in place as the reference's nagopt mutates it.
int orc_nagopt_cb(int prec, const long long *dims, orc_op_cb grad, orc_loss_cb loss, void *user, double *x3,
                  long long max_iters, double lipschitz, double tol, double xtol, long long max_inner) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            const size_t n = d.size();
            Vec3<T> x = load3<T>(x3, n);
            auto lf = [=](const Vec3<T> &v) {
                std::vector<double> in(3 * n);
                store3(v, in.data());
                return loss(user, in.data());
            };
            Vec3<T> r = nagopt<T>(cb_op<T>(grad, user, n), lf, x, d, (size_t)max_iters, (T)lipschitz, (T)tol,
                                  (T)xtol, (size_t)max_inner);
            store3(r, x3);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

int orc_lipschitz_cb(int prec, const long long *dims, orc_op_cb A, void *user, const double *x03,
                     long long max_iter, double tol, double *L) {
    return guard([&] {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Dims d = mkd(dims);
            *L = (double)lipschitz<T>(cb_op<T>(A, user, d.size()), load3<T>(x03, d.size()), d, (size_t)max_iter,
                                      (T)tol);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// The MBIR drivers' shared set-up: padded dims (od3) and yT [3][od].  Call
// with yT3 = nullptr to get the dims only.
int orc_mbir_setup(int prec, double eps, int method, const long long *recon_dims, int G, const long long *P,
                   const long long *R, const long long *C, const double *proj, const double *alpha,
                   const double *gamma, long long *od3, double *yT3) {
    return guard([&] {
        Params pr;
        pr.method = method;
        Ctx cx; cx.eps = eps;
        auto ds = mkds(G, P, R, C, proj, alpha, gamma);
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Job<T> job = mbir_setup<T>(cx, ds, mkd(recon_dims), pr);
            od3[0] = job.od.n1; od3[1] = job.od.n2; od3[2] = job.od.n3;
            if (yT3) store3(job.yT, yT3);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// MBIR3's gradient and loss closures at a padded iterate x3.
int orc_mbir3_closures(int prec, double eps, const long long *recon_dims, int G, const long long *P,
                       const long long *R, const long long *C, const double *proj, const double *alpha,
                       const double *gamma, double sigma, double p, const double *x3, double *g3, double *loss) {
    return guard([&] {
        Params pr;
        pr.method = 3; pr.sigma = sigma; pr.p = p;
        Ctx cx; cx.eps = eps;
        auto ds = mkds(G, P, R, C, proj, alpha, gamma);
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Job<T> job = mbir_setup<T>(cx, ds, mkd(recon_dims), pr);
            Vec3<T> x = load3<T>(x3, job.od.size());
            store3(mbir3_grad<T>(cx, job, x, pr), g3);
            *loss = mbir3_loss<T>(cx, job, x);
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

// MBIR1/2/3.  Datasets are concatenated: proj (sum_g P_g R_g C_g), alpha/gamma (sum_g P_g).
int orc_mbir(int prec, double eps, int method, const long long *recon_dims, int G, const long long *P,
             const long long *R, const long long *C, const double *proj, const double *alpha, const double *gamma,
             long long max_iters, long long inner_iters, double lambda, double mu, double sigma, double p, double tol,
             double xtol, double *out3) {
    return guard([&] {
        std::vector<Dataset> ds;
        size_t po = 0, ao = 0;
        for (int g = 0; g < G; ++g) {
            ds.push_back(Dataset{(size_t)P[g], (size_t)R[g], (size_t)C[g], proj + po, alpha + ao, gamma + ao});
            po += (size_t)(P[g] * R[g] * C[g]);
            ao += (size_t)P[g];
        }
        Params pr;
        pr.method = method; pr.max_iters = max_iters; pr.inner_iters = inner_iters;
        pr.lambda = lambda; pr.mu = mu; pr.sigma = sigma; pr.p = p; pr.tol = tol; pr.xtol = xtol;
        Ctx cx; cx.eps = eps;
        Dims rd = mkd(recon_dims);
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Vec3<T> out = mbir<T>(cx, ds, rd, pr);
            for (int i = 0; i < 3; ++i) fromvec(out[i], out3 + i * rd.size());
        };
        if (prec == 32) run(float{}); else run(double{});
    });
}

}  // extern "C"
