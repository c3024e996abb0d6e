// oracle/ref_driver.cpp -- C entry points into the parts of the reference that
// compile from their own sources in this container (built by oracle/Makefile
// into oracle/_ref/libref.so together with the reference's src/finitediff.cpp
// and src/qggmrf.cpp, read where they lie).  TEST INFRASTRUCTURE ONLY: used to
// pin the oracle and to write tests/golden fixtures; never shipped.
//
//   PolarGrid ctor            include/polar_grid.h:44-75
//   beam_dir_vector           include/projection.h:36-41
//   fftshift2 / ifftshift2    include/fftutils.h:84-103, 147-168
//   pad2d / crop3d            include/padding.h:93-131, 206-232
//   grad_u / divergence       src/finitediff.cpp:7-71
//   laplacian                 include/bregman.h:33-36
//   qggmrf                    src/qggmrf.cpp:59-95
#include <complex>
#include <cstring>
#include <vector>

#include "bregman.h"
#include "fftutils.h"
#include "padding.h"
#include "polar_grid.h"
#include "projection.h"

namespace tomocam::opt {
// declared in the reference's include/optimize.h, which cannot be included
// here (it pulls <format> through array_ops.h); defined in src/qggmrf.cpp
template <typename T> void qggmrf(Array<T> &g, const Array<T> &x, T sigma, T p);
}  // namespace tomocam::opt

using tomocam::Array;
using tomocam::dims_t;

namespace {
template <class T> Array<T> load(const double *a, dims_t d) {
    Array<T> r(d);
    for (size_t i = 0; i < d.size(); ++i) r[i] = (T)a[i];
    return r;
}
template <class T> void store(const Array<T> &r, double *out) {
    for (size_t i = 0; i < r.size(); ++i) out[i] = (double)r[i];
}
dims_t mk(const long long *d) { return dims_t((size_t)d[0], (size_t)d[1], (size_t)d[2]); }
}  // namespace

extern "C" {

int ref_grid_coords(int prec, const double *theta, long long P, long long R, long long C, double gamma,
                    double *x, double *y, double *z) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        std::vector<T> ang(theta, theta + P);
        tomocam::PolarGrid<T> g(ang, (size_t)R, (size_t)C, (T)gamma);
        store(g.x, x); store(g.y, y); store(g.z, z);
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

int ref_beam_dir(int prec, double gamma, double alpha, double *out3) {
    if (prec == 32) {
        auto d = tomocam::beam_dir_vector<float>((float)gamma, (float)alpha);
        for (int i = 0; i < 3; ++i) out3[i] = d[i];
    } else {
        auto d = tomocam::beam_dir_vector<double>(gamma, alpha);
        for (int i = 0; i < 3; ++i) out3[i] = d[i];
    }
    return 0;
}

// complex interleaved in/out
int ref_shift2(const long long *dims, const double *in, int inverse, double *out) {
    dims_t d = mk(dims);
    Array<std::complex<double>> a(d);
    for (size_t i = 0; i < d.size(); ++i) a[i] = std::complex<double>(in[2 * i], in[2 * i + 1]);
    auto b = inverse ? tomocam::fft::ifftshift2(a) : tomocam::fft::fftshift2(a);
    for (size_t i = 0; i < d.size(); ++i) { out[2 * i] = b[i].real(); out[2 * i + 1] = b[i].imag(); }
    return 0;
}

// pad2d symmetric; out must hold the padded size, out_dims receives it
int ref_pad2d(int prec, const long long *dims, const double *in, double factor, double *out, long long *out_dims) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        auto a = load<T>(in, mk(dims));
        auto b = tomocam::pad2d<T>(a, (T)factor, tomocam::PadType::SYMMETRIC);
        if (out) store(b, out);
        out_dims[0] = (long long)b.nslices(); out_dims[1] = (long long)b.nrows(); out_dims[2] = (long long)b.ncols();
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

int ref_crop3d(const long long *dims, const double *in, const long long *new_dims, double *out) {
    auto a = load<double>(in, mk(dims));
    auto b = tomocam::crop3d<double>(a, mk(new_dims), tomocam::PadType::SYMMETRIC);
    store(b, out);
    return 0;
}

int ref_grad_u(int prec, const long long *dims, const double *u, double *g3) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        dims_t d = mk(dims);
        auto g = tomocam::opt::grad_u<T>(load<T>(u, d));
        for (int i = 0; i < 3; ++i) store(g[i], g3 + i * d.size());
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

int ref_divergence(int prec, const long long *dims, const double *v3, double *out) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        dims_t d = mk(dims);
        std::array<Array<T>, 3> v = {load<T>(v3, d), load<T>(v3 + d.size(), d), load<T>(v3 + 2 * d.size(), d)};
        store(tomocam::opt::divergence<T>(v), out);
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

int ref_laplacian(int prec, const long long *dims, const double *u, double *out) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        store(tomocam::opt::laplacian<T>(load<T>(u, mk(dims))), out);
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

int ref_qggmrf(int prec, const long long *dims, double *g, const double *x, double sigma, double p) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        dims_t d = mk(dims);
        auto ga = load<T>(g, d);
        tomocam::opt::qggmrf<T>(ga, load<T>(x, d), (T)sigma, (T)p);
        store(ga, g);
    };
    if (prec == 32) run(float{}); else run(double{});
    return 0;
}

}  // extern "C"
