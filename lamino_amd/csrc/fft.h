// Batched 1-D FFT passes over one axis of a strided array, in LDS.
//
// A pass transforms a set of lines indexed by (o1, o2, i); element n of line
// (o1, o2, i) lives at  o1*s1 + o2*s2 + i*si + n*sn  (separate strides for the
// input and the output, in elements of their own type).  One workgroup owns
// L lines: consecutive i when ci > 1 (so that the loads of a row of the tile
// are contiguous), otherwise consecutive o2.  The transform is one of
//   * a mixed-radix Stockham autosort for 2,3,5-smooth n, each stage one
//     register-resident radix of 2..12 (composites 6, 8, 9, 10, 12 are
//     Cooley-Tukey splits with compile-time twiddles), as few stages as the
//     radices allow (360 = 9 * 8 * 5, 90 = 10 * 9);
//   * for n = S * Q with S 2,3,5-smooth and Q > 5 a prime whose Q - 1 is
//     2,3,5-smooth (305 = 5 * 61, 615 = 15 * 41, 7, 11, 13, ...): the
//     Good-Thomas prime-factor split into S x Q (no twiddles between them),
//     Stockham over S, and Rader's algorithm over Q, which turns a length-Q
//     DFT into a cyclic convolution of length Q - 1 done with two smooth
//     Stockham FFTs;
//   * otherwise Bluestein's chirp-z algorithm on a 2,3,5-smooth inner length
//     m >= 2n - 1;
// all entirely in LDS between one load and one store.
//
// Fused at the load:  rotation (fftshift), zero padding beyond nv, real,
// real-pair (re from in, im from in2) or complex-pair (in + i in2) input.
// Fused at the store: rotation (ifftshift), truncation to ns, scaling, real or
// real-pair output.
#pragma once
#include "common.h"

namespace lam {

enum { LD_COMPLEX = 0, LD_REAL = 1, LD_REAL_PAIR = 2, LD_COMPLEX_PAIR = 3 };
enum { ST_COMPLEX = 0, ST_REAL = 1, ST_REAL_PAIR = 2, ST_IMAG = 3 };

struct FftPlan {
    int n = 0, sign = -1, m = 0;
    bool blue = false;
    // prime-factor + Rader plan (n = S * Q): rad[0, nrad_s) factor S, rad[nrad_s, nrad) factor Q - 1;
    // qdirect: the Q-point transforms run as one direct (symmetric-pair) DFT stage instead
    // of Rader's convolution, rad[] then factors S only
    bool rader = false, qdirect = false;
    int S = 1, Q = 1, nrad_s = 0;
    int nrad = 0;
    int rad[24] = {};
    DevBuf chirp;   // n entries exp(sign pi i k^2 / n)        (Bluestein)
    DevBuf bhat;    // m entries FFT_m(conj chirp, wrapped) / m (Bluestein);
                    // Rader: Q - 1 entries FFT_{Q-1}(b) / (Q - 1), b_j = exp(sign 2 pi i g^j / Q)
    DevBuf pos;     // Rader / qdirect: 2n shorts, the LDS slot of input element k, then of output element k
    // per stage of the transform's stage list: twiddles exp(sign 2 pi i t k / (Ns R)),
    // t = 1..R-1, for k < Ns, stored [k][t - 1] from float2 offset tw_off[stage]
    DevBuf tw;
    int tw_off[24] = {};
};

struct AxisPass {
    const void *in = nullptr, *in2 = nullptr;
    void *out = nullptr, *out2 = nullptr;
    int ld_mode = LD_COMPLEX, st_mode = ST_COMPLEX;
    long long is1 = 0, is2 = 0, isi = 0, isn = 1;
    long long os1 = 0, os2 = 0, osi = 0, osn = 1;
    int c1 = 1, c2 = 1, ci = 1;
    int nv = -1, ns = -1;      // -1: full length; valid loads m < nv, stored outputs m < ns
    int nv2 = 0, ns2 = 0;      // additionally valid / stored: the last nv2 / ns2 positions (wrapped ranges)
    // windows: when *_len > 0, only positions m with (m - *_lo) mod n < *_len are loaded (the
    // rest read as zero) / stored
    int ld_lo = 0, ld_len = 0, st_lo = 0, st_len = 0;
    int rot_in = 0, rot_out = 0;
    // pair modes: lines with o1 >= c1_in2 have no second input (in2 reads as zero),
    // lines with o1 >= c1_out2 store no second output
    int c1_in2 = 1 << 30, c1_out2 = 1 << 30;
    float scale = 1.0f;
    // line sets with a hole: line o maps to o < split ? o : o + gap (wrapped mode ranges)
    int o1_split = 1 << 30, o1_gap = 0;
    int o2_split = 1 << 30, o2_gap = 0;
};

void run_axis(Ctx &cx, const FftPlan &plan, AxisPass p);

// Convenience: the reference's centred 2-D detector transform
//   ifftshift2(fft2|ifft2(fftshift2(in)))  (include/fft.h:34-114,
//   include/fftutils.h:84-168, as used in src/projection.cpp:73-75, 94-99)
// over P planes of R x C.  in_real: float input; out_real: store Re()*scale.
void det_fft2(Ctx &cx, const void *in, bool in_real, void *out, bool out_real, int P, int R, int C, int sign,
              float scale, float2 *tmp);

// The same transform for real projections, two planes per complex transform
// (odd R and C only, where the mirrored detector point (R-1-j, C-1-k) is the
// transform's negative frequency).  spec holds P planes of R x C.
//   det_fft2_real_pairs:  spec plane 2m <- T(proj[2m] + i proj[2m+1]); the
//       plane spectra are P2m(q) = (Z(q) + conj Z(-q)) / 2 and
//       P2m+1(q) = (Z(q) - conj Z(-q)) / 2i.
//   det_ifft2_real_pairs: proj[2m] + i proj[2m+1] <- scale T(spec[2m] + i spec[2m+1])
//       for Hermitian plane spectra; spec's even planes are overwritten.
void det_fft2_real_pairs(Ctx &cx, const float *proj, float2 *spec, int P, int R, int C, int sign);
void det_ifft2_real_pairs(Ctx &cx, float2 *spec, float *proj, int P, int R, int C, int sign, float scale);

// The Toeplitz normal operator's spectral step along the contiguous axis of the
// L-grid: spectrum rows (z, y) of za/zb (transformed along z and y already,
// nonzero at x < nvx) are transformed along x, multiplied by the 3x3 kernel
// spectrum K ([L1][L2][L3][3], see toeplitz.h) and transformed back; wa/wb get
// x < nvx of every row.  L3 must be 2,3,5-smooth.
void toep_spectral_x(Ctx &cx, const float2 *za, const float2 *zb, const float2 *K, float2 *wa, float2 *wb, int L1,
                     int L2, int L3, int nvx);

// Full in-place 3-D complex FFT of [A][B][Cn][nc] (nc interleaved components).
void fft3d(Ctx &cx, float2 *data, int A, int B, int Cn, int nc, int sign);

}  // namespace lam
