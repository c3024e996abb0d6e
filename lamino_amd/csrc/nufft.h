// Polar-grid geometry and the 3-D NUFFT (exponential-of-semicircle kernel,
// upsampling 2) specialised to the plane-per-projection point sets of the
// reference's PolarGrid (include/polar_grid.h:44-75).
#pragma once
#include <algorithm>

#include "common.h"
#include "fft.h"

namespace lam {

// Per-projection plane: point (j,k) of projection p sits at
//   q = qX_k u_p + qY_j v_p,  qX_k = (k-(C-1)/2) 2pi/C,  qY_j = (j-(R-1)/2) 2pi/R
// with u_p = (cos g, 0, -sin g), v_p = (cos a sin g, -sin a, cos a cos g) in
// (x, y, z) = (n3, n2, n1) axis order, and beam direction
// d_p = (sin g sin a, cos a, cos g sin a) (include/projection.h:36-41).
struct Geo {
    double ux, uz, vx, vy, vz;
    double d0, d1, d2;
};

// Points of a grid bucketed by the fine-grid tile that holds their first
// kernel cell: ids[] lists point numbers tile by tile; chunks[] are
// (tile, first, count) work items of bounded size for the tiled spreader.
struct SortedPoints {
    DevBuf ids;
    DevBuf chunks;
    long long nchunks = 0;
    long long npoints = 0;
};

struct Grid {
    Ctx *cx = nullptr;
    long long P = 0, R = 0, C = 0;
    std::vector<double> alpha, gamma;
    bool literal = false;      // built by the reference ctor (constant gamma)
    double gamma_const = 0;
    DevBuf geo;                // Geo[P]
    std::vector<Geo> geo_host;
    std::map<std::vector<long long>, std::unique_ptr<SortedPoints>> bins;
    ~Grid();
};

struct NufftPlan {
    long long n[3] = {0, 0, 0};   // volume dims n1 (z), n2 (y), n3 (x)
    int nf[3] = {0, 0, 0};        // fine grid z, y, x
    int w = 6;
    float beta = 13.8f;
    DevBuf corr[3];               // float 1/psi_hat per axis, indexed by array index
    DevBuf fine;                  // float2 [nfz][nfy][nfx][3]
};

int width_for_eps(double eps);
// fine-grid length of one axis of n modes at tolerance eps (the plan's rule)
size_t nufft_fine_len(long long n, double eps);
NufftPlan &get_nufft_plan(Ctx &cx, const long long dims[3], double eps);
std::unique_ptr<Grid> make_grid(Ctx &cx, const double *alpha, const double *gamma, long long P, long long R,
                                long long C);

// Projectors and NUFFT entry points (device pointers, context stream)
void nufft2_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, float2 *c, const float2 *f);
void nufft1_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, const float2 *c, float2 *f);
void forward(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const m[3], float *proj);
void adjoint(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *proj, float *const m[3],
             bool accumulate);
void sysmat_nufft(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const x[3],
                  float *const y[3]);
// Picks the spreading/interpolation tile shape of this grid and fine grid by
// timing candidates; the choice is kept in cx.bin_tune.  best_ms: spread+interp time.
std::array<int, 3> autotune_bins(Ctx &cx, Grid &g, const long long dims[3], double eps, double *best_ms);
// whether a tuned shape exists (in this context or the process-wide table)
bool bins_tuned(Ctx &cx, Grid &g, const long long dims[3], double eps);

}  // namespace lam

struct lam_grid {
    std::unique_ptr<lam::Grid> impl;
};
