// Parity tests of the C++ drop-in layer (include/tomocam/*.h): the reference's
// own API (tomocam::forward, adjoint, gradient, residual, nufft::nufft3d1/2,
// opt::cgsolver, RampPreconditioner, grad_u/divergence/laplacian, qggmrf,
// MBIR1/2/3) called exactly as the reference's tests call it
// (tests/test_grad.cpp, tests/test_loss.cpp, tests/test_cg.cpp), on the GPU,
// compared with the oracle's FP64 restatement (oracle/liboracle.so).
//
// Usage: test_dropin [name...]   (no names: run every test).  Exit status 0
// when every selected test passes.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "tomocam/tomocam.h"

extern "C" {
int orc_forward(int prec, double eps, int fft_threads, const long long *dims, const double *m, const double *alpha,
                const double *gamma, long long P, long long R, long long C, double *proj);
int orc_adjoint(int prec, double eps, int fft_threads, const long long *dims, const double *proj,
                const double *alpha, const double *gamma, long long P, long long R, long long C, double *m);
int orc_sysmat(int prec, double eps, const long long *dims, const double *x, int G, const long long *P,
               const long long *R, const long long *C, const double *alpha, const double *gamma, double *y);
int orc_nufft(int prec, int type, int brute, double eps, const long long *dims, const double *alpha,
              const double *gamma, long long P, long long R, long long C, const double *in, double *out);
int orc_grid_coords(int prec, const double *alpha, const double *gamma, long long P, long long R, long long C,
                    double *x, double *y, double *z);
int orc_precond(int prec, const long long *dims, const double *in, double *out);
int orc_grad_u(int prec, const long long *dims, const double *u, double *g3);
int orc_divergence(int prec, const long long *dims, const double *v3, double *out);
int orc_qggmrf(int prec, const long long *dims, double *g, const double *x, double sigma, double p);
int orc_cg_diag(int prec, const long long *dims, const double *diag3, const double *b3, const double *x03,
                long long max_iter, double tol, double *x3, int *iters);
int orc_mbir(int prec, double eps, int method, const long long *recon_dims, int G, const long long *P,
             const long long *R, const long long *C, const double *proj, const double *alpha, const double *gamma,
             long long max_iters, long long inner_iters, double lambda, double mu, double sigma, double p, double tol,
             double xtol, double *out3);
const char *orc_last_error();
}

using namespace tomocam;
using V3 = std::array<Array<float>, 3>;

namespace {

constexpr double kFwdTol = 1e-4;   // forward / adjoint / NUFFT vs the FP64 oracle
constexpr double kRecTol = 1e-3;   // reconstructions vs the FP64 oracle
constexpr double kOracleEps = 1e-12;

int g_fail = 0;

void orc(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("oracle: ") + orc_last_error());
}

void report(const char *name, const char *what, double value, double tol) {
    const bool ok = value < tol;
    std::printf("%-28s %-26s %.3e (< %.0e) %s\n", name, what, value, tol, ok ? "PASS" : "FAIL");
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

void expect(const char *name, const char *what, bool ok) {
    std::printf("%-28s %-26s %s\n", name, what, ok ? "PASS" : "FAIL");
    std::fflush(stdout);
    if (!ok) ++g_fail;
}

// SplitMix64 uniform [-1, 1)
struct Rng {
    uint64_t s;
    double next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-52 - 1.0;
    }
};

Array<float> rand_array(dims_t d, Rng &r) {
    Array<float> a(d);
    for (auto &v : a) v = (float)r.next();
    return a;
}

V3 rand_vec(dims_t d, Rng &r) { return {rand_array(d, r), rand_array(d, r), rand_array(d, r)}; }

template <typename A>
std::vector<double> flat(const A &a) { return std::vector<double>(a.begin(), a.end()); }

std::vector<double> flat3(const V3 &v) {
    std::vector<double> o;
    for (const auto &c : v) o.insert(o.end(), c.begin(), c.end());
    return o;
}

double rel(const std::vector<double> &a, const std::vector<double> &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1.0));
}

long long *ll3(dims_t d, long long *o) {
    o[0] = (long long)d.n1; o[1] = (long long)d.n2; o[2] = (long long)d.n3;
    return o;
}

// tests/test_grad.cpp geometry: theta_i = (i - 70) deg around 0, gamma = pi/4
struct Scan {
    std::vector<float> theta;
    float gamma = 0.7853981634f;
    size_t R, C;
    std::vector<double> alpha_d, gamma_d;
    Scan(size_t n, size_t R_, size_t C_) : R(R_), C(C_) {
        for (size_t i = 0; i < n; ++i) theta.push_back((float)(((double)i - (double)(n - 1)) * M_PI / 180.0 * 5.0));
        alpha_d.assign(theta.begin(), theta.end());
        gamma_d.assign(n, (double)gamma);
    }
    size_t P() const { return theta.size(); }
};

void test_grid_coords() {
    const char *N = "grid_coords";
    Scan s(7, 9, 11);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    const size_t n = pg.size();
    std::vector<double> x(n), y(n), z(n);
    orc(orc_grid_coords(32, s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C, x.data(), y.data(), z.data()));
    double m = 0;
    for (size_t k = 0; k < n; ++k)
        m = std::max({m, std::abs(pg.x[k] - x[k]), std::abs(pg.y[k] - y[k]), std::abs(pg.z[k] - z[k])});
    report(N, "max |dx| (fp32 grid)", m, 1e-5);
    expect(N, "dims (P, R, C)", pg.x.dims() == dims_t(s.P(), s.R, s.C) && pg.npts == n);
}

void test_forward_adjoint() {
    const char *N = "forward_adjoint";
    Rng r{101};
    dims_t d(9, 5, 11);
    Scan s(12, 9, 11);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    V3 f = rand_vec(d, r);
    Array<float> p = forward(f, pg, s.gamma);
    long long dd[3];
    std::vector<double> want(pg.size());
    auto fm = flat3(f);
    orc(orc_forward(64, kOracleEps, 0, ll3(d, dd), fm.data(), s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C,
                    want.data()));
    report(N, "forward rel", rel(flat(p), want), kFwdTol);
    auto a = adjoint(p, pg, d, s.gamma);
    std::vector<double> pm = flat(p), wa(3 * d.size());
    orc(orc_adjoint(64, kOracleEps, 0, dd, pm.data(), s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C,
                    wa.data()));
    report(N, "adjoint rel", rel(flat3(a), wa), kFwdTol);
    // <A f, p> = <f, A^T p>
    double l = 0, rr = 0;
    for (size_t k = 0; k < p.size(); ++k) l += (double)p[k] * p[k];
    for (int i = 0; i < 3; ++i)
        for (size_t k = 0; k < d.size(); ++k) rr += (double)f[i][k] * a[i][k];
    report(N, "adjointness", std::abs(l - rr) / std::abs(l), 1e-5);
}

// tests/test_grad.cpp: gradient(f, 0) == adjoint(forward(f)) == sysmat(f)
void test_grad() {
    const char *N = "gradient";
    Rng r{202};
    dims_t d(9, 5, 11);
    Scan s(15, 9, 11);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    V3 f = rand_vec(d, r);
    auto direct = adjoint(forward(f, pg, s.gamma), pg, d, s.gamma);
    V3 yT{Array<float>::zeros(d), Array<float>::zeros(d), Array<float>::zeros(d)};
    auto g = gradient(f, yT, pg, s.gamma);
    report(N, "adjoint(forward) vs grad", rel(flat3(g), flat3(direct)), kFwdTol);
    long long dd[3], P = s.P(), R = s.R, C = s.C;
    std::vector<double> fm = flat3(f), want(3 * d.size());
    orc(orc_sysmat(64, kOracleEps, ll3(d, dd), fm.data(), 1, &P, &R, &C, s.alpha_d.data(), s.gamma_d.data(),
                   want.data()));
    report(N, "sysmat vs oracle", rel(flat3(sysmat(f, pg, s.gamma)), want), kFwdTol);
}

// tests/test_loss.cpp: residual from the normal operator
void test_loss() {
    const char *N = "residual";
    Rng r{303};
    dims_t d(7, 5, 9);
    Scan s(10, 9, 9);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    V3 f = rand_vec(d, r);
    Array<float> y = rand_array(pg.dims(), r);
    const float yTy = array::dot(y, y);
    auto yT = adjoint(y, pg, d, s.gamma);
    const float res = residual(f, yT, pg, yTy, s.gamma);
    long long dd[3], P = s.P(), R = s.R, C = s.C;
    std::vector<double> fm = flat3(f), Af(3 * d.size());
    orc(orc_sysmat(64, kOracleEps, ll3(d, dd), fm.data(), 1, &P, &R, &C, s.alpha_d.data(), s.gamma_d.data(),
                   Af.data()));
    double want = 0;
    for (int i = 0; i < 3; ++i) {
        double xAx = 0, yx = 0;
        for (size_t k = 0; k < d.size(); ++k) {
            xAx += fm[i * d.size() + k] * Af[i * d.size() + k];
            yx += fm[i * d.size() + k] * (double)yT[i][k];
        }
        want += std::sqrt(xAx - 2 * yx + (double)yTy);
    }
    report(N, "residual rel", std::abs(res - want) / std::abs(want), kFwdTol);
}

void test_nufft() {
    const char *N = "nufft3d1_3d2";
    Rng r{404};
    dims_t d(5, 3, 7);
    Scan s(5, 7, 9);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    Array<std::complex<float>> fz(d), cz(pg.dims());
    for (auto &v : fz) v = {(float)r.next(), (float)r.next()};
    nufft::nufft3d2(cz, fz, pg);
    long long dd[3];
    std::vector<double> in(2 * fz.size()), want(2 * cz.size());
    for (size_t k = 0; k < fz.size(); ++k) { in[2 * k] = fz[k].real(); in[2 * k + 1] = fz[k].imag(); }
    orc(orc_nufft(64, 2, 1, 0.0, ll3(d, dd), s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C, in.data(),
                  want.data()));
    std::vector<double> got(2 * cz.size());
    for (size_t k = 0; k < cz.size(); ++k) { got[2 * k] = cz[k].real(); got[2 * k + 1] = cz[k].imag(); }
    report(N, "type 2 vs NUDFT", rel(got, want), kFwdTol);

    Array<std::complex<float>> c1(pg.dims()), f1(d);
    for (auto &v : c1) v = {(float)r.next(), (float)r.next()};
    nufft::nufft3d1(c1, f1, pg);
    std::vector<double> in1(2 * c1.size()), want1(2 * f1.size()), got1(2 * f1.size());
    for (size_t k = 0; k < c1.size(); ++k) { in1[2 * k] = c1[k].real(); in1[2 * k + 1] = c1[k].imag(); }
    orc(orc_nufft(64, 1, 1, 0.0, dd, s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C, in1.data(),
                  want1.data()));
    for (size_t k = 0; k < f1.size(); ++k) { got1[2 * k] = f1[k].real(); got1[2 * k + 1] = f1[k].imag(); }
    report(N, "type 1 vs NUDFT", rel(got1, want1), kFwdTol);
}

void test_stencils() {
    const char *N = "precond_stencils_qggmrf";
    Rng r{505};
    dims_t d(6, 7, 9);
    long long dd[3];
    ll3(d, dd);
    Array<float> u = rand_array(d, r);
    std::vector<double> um = flat(u);

    opt::RampPreconditioner<float> pc(d);
    std::vector<double> w(d.size());
    orc(orc_precond(64, dd, um.data(), w.data()));
    report(N, "precond rel", rel(flat(pc.apply(u)), w), 5e-6);

    std::vector<double> g3(3 * d.size());
    orc(orc_grad_u(64, dd, um.data(), g3.data()));
    auto g = opt::grad_u(u);
    report(N, "grad_u rel", rel(flat3(g), g3), 1e-6);

    std::vector<double> dv(d.size());
    orc(orc_divergence(64, dd, g3.data(), dv.data()));
    report(N, "divergence rel", rel(flat(opt::divergence(g)), dv), 1e-6);
    report(N, "laplacian rel", rel(flat(opt::laplacian(u)), dv), 1e-6);

    Array<float> gq = rand_array(d, r);
    std::vector<double> gw = flat(gq);
    orc(orc_qggmrf(64, dd, gw.data(), um.data(), 1.5, 1.2));
    opt::qggmrf(gq, u, 1.5f, 1.2f);
    report(N, "qggmrf rel", rel(flat(gq), gw), 1e-5);
}

// tests/test_cg.cpp: CG through a host std::function operator
void test_cg() {
    const char *N = "cgsolver";
    Rng r{606};
    dims_t d(4, 5, 6);
    V3 diag, b, x0;
    for (int i = 0; i < 3; ++i) {
        diag[i] = Array<float>(d);
        for (size_t k = 0; k < d.size(); ++k) diag[i][k] = 1.0f + (float)(i * d.size() + k) / 10.0f;
        b[i] = rand_array(d, r);
        x0[i] = Array<float>::zeros(d);
    }
    opt::Function<float> A = [&](const V3 &v) {
        V3 o{Array<float>(d), Array<float>(d), Array<float>(d)};
        for (int i = 0; i < 3; ++i)
            for (size_t k = 0; k < d.size(); ++k) o[i][k] = diag[i][k] * v[i][k];
        return o;
    };
    auto x = opt::cgsolver(A, b, x0, 25, 0.0f, 0.0f);
    long long dd[3];
    std::vector<double> dg = flat3(diag), bm = flat3(b), xm = flat3(x0), want(3 * d.size());
    int it = 0;
    orc(orc_cg_diag(64, ll3(d, dd), dg.data(), bm.data(), xm.data(), 25, 0.0, want.data(), &it));
    report(N, "x rel", rel(flat3(x), want), 1e-4);
}

void test_mbir() {
    const char *N = "MBIR1_MBIR2";
    Rng r{707};
    dims_t rd(9, 5, 11);
    std::vector<std::tuple<Array<float>, std::vector<float>, float>> ds;
    std::vector<double> projm, alpham, gammam;
    std::vector<long long> Ps, Rs, Cs;
    for (int g = 0; g < 2; ++g) {
        const size_t P = 8, det = 9;
        Array<float> p(P, det, det);
        for (auto &v : p) v = (float)(3.0 + r.next());
        std::vector<float> ang;
        for (size_t j = 0; j < P; ++j) ang.push_back((float)(-1.2 + 1.2 * (double)j / (P - 1)));
        const float gamma = (float)(M_PI / 4 * g);
        projm.insert(projm.end(), p.begin(), p.end());
        for (float a : ang) { alpham.push_back(a); gammam.push_back(gamma); }
        Ps.push_back(P); Rs.push_back(det); Cs.push_back(det);
        ds.emplace_back(std::move(p), ang, gamma);
    }
    long long rdd[3];
    ll3(rd, rdd);
    ReconParams prm;
    prm.maxIters = 6;
    prm.tol = 0.0f;
    std::vector<double> want(3 * rd.size());
    // MBIR1 on the first dataset
    std::vector<std::tuple<Array<float>, std::vector<float>, float>> ds1;
    ds1.emplace_back(std::get<0>(ds[0]).clone(), std::get<1>(ds[0]), std::get<2>(ds[0]));
    orc(orc_mbir(64, kOracleEps, 1, rdd, 1, Ps.data(), Rs.data(), Cs.data(), projm.data(), alpham.data(),
                 gammam.data(), 6, 3, 0.1, 10.0, 1000.0, 1.2, 0.0, 1e-5, want.data()));
    report(N, "MBIR1 rel", rel(flat3(MBIR1(ds1, rd, prm)), want), kRecTol);
    // MBIR2 on both
    ReconParams p2;
    p2.maxIters = 3;
    p2.innerIters = 2;
    p2.tol = 0.0f;
    p2.xtol = 0.0f;
    orc(orc_mbir(64, kOracleEps, 2, rdd, 2, Ps.data(), Rs.data(), Cs.data(), projm.data(), alpham.data(),
                 gammam.data(), 3, 2, 0.1, 10.0, 1000.0, 1.2, 0.0, 0.0, want.data()));
    report(N, "MBIR2 rel", rel(flat3(MBIR2(ds, rd, p2)), want), kRecTol);
}

// Exceptions the reference throws for the same misuse.
void test_errors() {
    const char *N = "errors";
    dims_t d(4, 5, 6), d2(4, 5, 7);
    V3 x0{Array<float>::zeros(d), Array<float>::zeros(d), Array<float>::zeros(d)};
    V3 yT{Array<float>::zeros(d2), Array<float>::zeros(d2), Array<float>::zeros(d2)};
    opt::Function<float> I = [](const V3 &v) { return V3{v[0].clone(), v[1].clone(), v[2].clone()}; };
    bool ok = false;
    try {
        opt::split_bregman(I, yT, x0, 0.1f, 10.0f, 2, 2, 0.0f, 0.0f);
    } catch (const std::invalid_argument &e) {
        ok = std::strstr(e.what(), "same dimensions as x0") != nullptr;
    }
    expect(N, "split_bregman shape", ok);

    ok = false;
    try {
        Array<float> a = Array<float>::ones(d);
        a /= 0.0f;
    } catch (const std::runtime_error &e) {
        ok = std::string(e.what()) == "Division by zero in Array<T>::operator/=";
    }
    expect(N, "Array /= 0", ok);

    Scan s(3, 5, 5);
    PolarGrid<float> pg(s.theta, s.R, s.C, s.gamma);
    ok = false;
    try {
        Array<std::complex<float>> fz(d), cz(2, 2, 2);
        nufft::nufft3d2(cz, fz, pg);
    } catch (const std::invalid_argument &) {
        ok = true;
    }
    expect(N, "nufft size mismatch", ok);

    // an exception thrown inside a host operator surfaces from the solver
    ok = false;
    opt::Function<float> bad = [](const V3 &) -> V3 { throw std::domain_error("operator failed"); };
    try {
        opt::cgsolver(bad, x0, x0, 3, 0.0f, 0.0f);
    } catch (const std::domain_error &) {
        ok = true;
    }
    expect(N, "operator exception", ok);
}

// The double instantiation (the reference instantiates both) computes in FP32 on the device.
void test_double() {
    const char *N = "double_api";
    Rng r{808};
    dims_t d(7, 5, 9);
    Scan s(8, 9, 9);
    std::vector<double> th(s.theta.begin(), s.theta.end());
    PolarGrid<double> pg(th, s.R, s.C, (double)s.gamma);
    std::array<Array<double>, 3> f{Array<double>(d), Array<double>(d), Array<double>(d)};
    for (auto &c : f)
        for (auto &v : c) v = r.next();
    // without the opt-in the double API refuses to compute in FP32
    unsetenv("LAMINO_FP32_FOR_DOUBLE");
    bool refused = false;
    try {
        (void)forward(f, pg, (double)s.gamma);
    } catch (const std::invalid_argument &e) {
        refused = std::strstr(e.what(), "LAMINO_FP32_FOR_DOUBLE") != nullptr;
    }
    report(N, "double refused without opt-in", refused ? 0.0 : 1.0, 0.5);
    setenv("LAMINO_FP32_FOR_DOUBLE", "1", 1);
    Array<double> p = forward(f, pg, (double)s.gamma);
    long long dd[3];
    std::vector<double> fm, want(pg.size());
    for (auto &c : f) fm.insert(fm.end(), c.begin(), c.end());
    orc(orc_forward(64, kOracleEps, 0, ll3(d, dd), fm.data(), s.alpha_d.data(), s.gamma_d.data(), s.P(), s.R, s.C,
                    want.data()));
    report(N, "forward<double> rel", rel(flat(p), want), kFwdTol);
}

struct Case {
    const char *name;
    void (*fn)();
};

}  // namespace

int main(int argc, char **argv) {
    const Case cases[] = {{"grid_coords", test_grid_coords}, {"forward_adjoint", test_forward_adjoint},
                          {"gradient", test_grad},           {"residual", test_loss},
                          {"nufft", test_nufft},             {"stencils", test_stencils},
                          {"cgsolver", test_cg},             {"mbir", test_mbir},
                          {"errors", test_errors},           {"double", test_double}};
    for (const Case &c : cases) {
        bool run = argc < 2;
        for (int a = 1; a < argc; ++a) run |= std::strcmp(argv[a], c.name) == 0;
        if (!run) continue;
        std::printf("== %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception &e) {
            std::printf("%-28s exception: %s FAIL\n", c.name, e.what());
            ++g_fail;
        }
    }
    std::printf("## %s (%d failed)\n", g_fail ? "FAILED" : "ALL PASSED", g_fail);
    return g_fail ? 1 : 0;
}
