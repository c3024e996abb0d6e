// Polar-grid geometry and the 3-D NUFFT (exponential-of-semicircle kernel,
// upsampling 2) specialised to the plane-per-projection point sets of the
// reference's PolarGrid (include/polar_grid.h:44-75).
#pragma once
#include <algorithm>

#include "common.h"
#include "fastdiv.h"
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

// Row-factorised form of the three-component projector NUFFT, for grids whose
// projections share the tilt alpha: then v_p has the same y component for
// every p and u_p has none, so a point's y = qY_j (-sin alpha) depends on its
// detector row j only.  The ES kernel factorises as wy_j(y) x [wz wx](z, x):
//   type 2: G_j(z, x) = sum_b wy_j[b] F(z, iy_j + b, x), then a 2-D gather of
//           W^2 taps per point from the (z, x) slab of its row;
//   type 1: a 2-D spread of W^2 taps per point into H_j(z, x), then
//           F(z, iy_j + b, x) += wy_j[b] H_j(z, x).
// The (z, x) plane is cut into tiles whose edges divide the fine-grid
// lengths; one workgroup owns one (row, tile).  The spreader stores its
// tile's interior to the slab and the parts of its box that reach into the
// next tiles (south and east strips of W - 1 cells, and their corner) to a
// halo buffer, and the y pass adds the halos of a cell's neighbours, so no
// value is written twice and no atomics are needed.  Rows are handled in
// groups whose slabs and halos fit a byte budget.
struct RowGeom {
    int Tz = 0, Tx = 0;        // tile edges, dividing nfz and nfx
    int ntz = 0, ntx = 0;      // tiles per axis
    unsigned ntiles = 0;
    int HBz = 0, HBx = 0;      // box extents Tz + W - 1, Tx + W - 1
    int NH = 0;                // halo cells per tile: (W-1) Tx + Tz (W-1) + (W-1)^2
    int XB = 1;                // spreader x block (divides Tx)
    FastDiv fbx;               // HBx
    FastDiv fbx3;              // 3 * HBx: complex values per box row of 3 components
};

struct RowPlan {
    RowGeom rg;
    int Jn = 0;                        // rows of the half point set
    int dir = 1;                       // +1: iy non-decreasing in j, -1: non-increasing
    int ylo = 0, ly = 0;               // the y planes the rows touch: [ylo, ylo + ly) mod nfy (ly 0: all)
    DevBuf wy;                         // float [Jn][W] row weights
    DevBuf iy;                         // int [Jn] unwrapped first y cell of each row
    std::vector<int> iy_host;          // the same on the host
    DevBuf count;                      // unsigned [Jn][ntiles] points per (row, tile)
    DevBuf nz;                         // unsigned [ntiles][nzw]: bit j % 32 of word j / 32 set when (row j, tile) holds points
    int nzw = 0;
    DevBuf ids;                        // half-point ids sorted by (row, tile, first cell)
    DevBuf chunks;                     // uint4 (row * ntiles + tile, first, count, 0), one per non-empty (row, tile)
    long long nchunks = 0;
    std::vector<long long> row_chunk;  // first chunk of each row, then nchunks (Jn + 1 entries)
};

struct Grid {
    Ctx *cx = nullptr;
    long long P = 0, R = 0, C = 0;
    std::vector<double> alpha, gamma;
    bool literal = false;      // built by the reference ctor (constant gamma)
    bool shared_tilt = false;  // every projection has the same alpha (row-factorised projectors)
    double gamma_const = 0;
    DevBuf geo;                // Geo[P]
    std::vector<Geo> geo_host;
    std::map<std::vector<long long>, std::unique_ptr<SortedPoints>> bins;
    std::map<std::vector<long long>, std::unique_ptr<RowPlan>> rowplans;
    ~Grid();
};

struct NufftPlan {
    long long n[3] = {0, 0, 0};   // volume dims n1 (z), n2 (y), n3 (x)
    int nf[3] = {0, 0, 0};        // fine grid z, y, x
    int w = 6;
    float beta = 13.8f;
    DevBuf corr[3];               // float 1/psi_hat per axis, indexed by array index
    DevBuf fine;                  // float2 [nfz][nfy][nfx][3]
    DevBuf slab;                  // float2 [rows][nfz][nfx][3] of the row-factorised path
    DevBuf halo;                  // float2 [rows][tiles][halo cells][3] of the row-factorised spreader
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
// Several grids at once: one fine-grid transform (and de-apodisation) serves them all.
// forward_multi: proj[k] = forward(m, grid k).  adjoint_multi: m (+)= sum_k adjoint(proj[k],
// grid k); the grids share the detector size.
void forward_multi(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], double eps,
                   const float *const m[3], float *const *proj);
void adjoint_multi(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], double eps,
                   const float *const *proj, float *const m[3], bool accumulate);
// Picks the spreading/interpolation tile shape of this grid and fine grid by
// timing candidates; the choice is kept in cx.bin_tune.  best_ms: spread+interp time.
std::array<int, 3> autotune_bins(Ctx &cx, Grid &g, const long long dims[3], double eps, double *best_ms);
// whether a tuned shape exists (in this context or the process-wide table)
bool bins_tuned(Ctx &cx, Grid &g, const long long dims[3], double eps);
// Device bytes of the row-factorised projectors of one shared-tilt grid of P
// projections of R x C on the fine grid of dims: slab: the row group's slabs
// and halos (kept by the plan, shared by grids of the same dims), kept: the
// grid's row bins, setup: temporaries while they are built.  Zero when the
// fine grid has no row tiling.
struct RowBytes {
    size_t slab = 0, kept = 0, setup = 0;
};
RowBytes row_path_bytes(const long long dims[3], double eps, long long P, long long R, long long C);

}  // namespace lam

struct lam_grid {
    std::unique_ptr<lam::Grid> impl;
};
