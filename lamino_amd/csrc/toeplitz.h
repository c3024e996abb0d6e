// Toeplitz form of the normal operator sum_g sysmat_g (src/gradient.cpp:33-89).
//
// sysmat_g(x)_i[m'] = sum_k sum_m T^g_ik(m' - m) x_k[m] with the exact kernel
//   T^g_ik(D) = 1/(R C) sum_p d_i(p) d_k(p) Dir_C(2pi/C D.u_p) Dir_R(2pi/R D.v_p),
//   Dir_n(t) = sin(n t / 2) / sin(t / 2),  D in (-n, n)^3,
// so the operator is a 3x3 block convolution evaluated with zero-padded FFTs on
// an L-grid (L_a >= 2 n_a - 1, 2,3,5-smooth).  The kernels are real and even,
// hence their spectra are real: pairs of them are stored as one complex array.
#pragma once
#include "common.h"
#include "nufft.h"
#include "ops.h"

namespace lam {

struct Toep {
    Ctx *cx = nullptr;
    long long n[3] = {0, 0, 0};
    int L[3] = {0, 0, 0};
    int ngrids = 0;
    bool per_grid = false;
    DevBuf khat;                    // float2 [L1][L2][L3][3]: (K00,K01),(K02,K11),(K12,K22) spectra / Ltot
    std::vector<DevBuf> khat_g;     // per-grid spectra (when per_grid)
    DevBuf za, zb, wa, wb;          // work spectra (complex L-grid each; wa, wb only without fusion)
    bool fused = true;              // x transforms + multiply in one pass (toep_spectral_x)
    long long applies = 0;
};

// L-grid length of one axis of n voxels: the smallest 2,3,5-smooth L >= 2n - 1
int toep_len(long long n);

std::unique_ptr<Toep> toep_create(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], bool per_grid);
// y = sum_g sysmat_g(x) + mu * laplacian(x)
void toep_apply(Toep &t, CVol3 x, Vol3 y, float mu);
// q[g*3+i] = x_i . sysmat_g(x)_i  (device FP64 array of 3*ngrids; no sync).  partials:
// device doubles for the block sums (toep_quad_partials() of them), or null for the
// context's shared reduction scratch (whose address may change; graphs pass their own)
void toep_quad_dev(Toep &t, CVol3 x, double *q_dev, double *partials = nullptr);
size_t toep_quad_partials();
void toep_quad(Toep &t, CVol3 x, double *q_host);

}  // namespace lam

struct lam_toep {
    std::unique_ptr<lam::Toep> impl;
};
