// LDS-resident mixed-radix Stockham / Bluestein FFT passes for gfx950.
#include <cmath>
#include <algorithm>
#include <complex>
#include <cstdlib>
#include <vector>

#include "fastdiv.h"
#include "fft.h"

namespace lam {

namespace {

// Workgroup size; with register radices up to 12 a stage keeps most of the 256
// threads busy (512 measured no faster).
constexpr int kThreads = 256;
// The workgroup's data tile: L * m <= 3072 complex (24 KiB) plus its line-base
// tables (2 x L x 8 B) and a Rader plan's slot tables.  Small tiles with several
// workgroups per CU hide the load/stage/store latency better than fewer, longer
// ones; with kFftWaves below, 6 workgroups of 4 waves share a CU (measured: 0.074
// ms average pass at 30 KiB and 5 waves, 0.068 here, 0.070 at 20 KiB and 7 waves,
// 0.079 at 40 KiB, 0.113 at 60 KiB).
constexpr int kLdsComplex = 3072;
// Longest transform.  Lines longer than kLdsComplex run one per workgroup in a
// second instantiation of the pass whose per-thread butterfly slots are sized
// for kMaxLen (and that carries no occupancy hint: it would spill).
constexpr int kMaxLen = 3840;
// Occupancy hint (waves per SIMD): the pass fits in 80 VGPRs without spills at
// this tile size and kLoadBatch, instead of 95.
constexpr int kFftWaves = 6;
constexpr int kMaxLines = 128;
// Longest 4 lines of the fused Toeplitz spectral step (lines up to 1536 points).
constexpr int kFusedMaxLen = 6144;

// Elements per thread whose global loads (load phase) or LDS reads (store
// phase) are issued together (8 measured the same and needs more registers).
constexpr int kLoadBatch = 4;

// Lines per tile along the inner (contiguous) index, LDS permitting: measured
// 16 -> 48 took the 90-point passes of the benchmark grid from 94 to 76 us.
constexpr int kInnerLines = 48;

// A set of nv equal sub-lines of the tile (prime-factor plans): sub-line v is
// column c of tile line l (INNER: v = c * L + l, else v = l * C + c), and its
// element t sits at position off + c * coff + t * str of line l.
struct SubLines {
    int nv = 0, C = 1, off = 0, coff = 0, str = 1;
    FastDiv fnv, fC;
};

enum { KIND_STOCKHAM = 0, KIND_BLUESTEIN = 1, KIND_RADER = 2 };

// Every index split by a launch-time divisor goes through FastDiv: a plain
// integer division costs ~30 VALU instructions and the passes do several per
// element.
//
// The transform is a list of Stockham stages, each over a (sub-)line set:
// st_g = 0: the tile's L lines of length m; 1: the S-point columns and 2: the
// (Q - 1)-point rows of a Rader plan.  Bluestein and Rader plans apply
// their pointwise product after stage pw_after.
struct KParams {
    AxisPass a;
    int n, m, kind, sign;
    int L, Li, Lo;  // lines per workgroup, L = Li * Lo
    int nst, pw_after;
    int rad[24], st_len[24], st_sign[24], st_g[24], st_ns[24], st_tw[24];
    const float2 *tw;                     // the plan's twiddle tables (stage s from st_tw[s])
    const float2 *chirp, *bhat;
    FastDiv fL, fN, fM, fLi, fPad;        // tile lines, length, inner length, lines per o2 row, M - N
    FastDiv fNbf[24], fNs[24];            // per stage: butterflies per line (len / R), span Ns
    int isn, osn;                         // element strides along the transform (fit 32 bits)
    // Rader plans: n = S * Q; the S-point transforms run over the Q columns of
    // each line (sub_s), the Q - 1 point convolutions over its S rows (sub_q)
    int S, Q;
    const short *pos_in, *pos_out;
    SubLines whole, sub_s, sub_q;
    FastDiv fQ1, fS;                      // Q - 1, S
};

// LDS position of element n of tile line l: n-major when the tile runs along
// the inner (contiguous) index, line-major otherwise.
template <int INNER>
__device__ __forceinline__ int lidx(int l, int n, int L, int M) {
    return INNER ? n * L + l : l * M + n;
}

// Element e of a stage over nv (sub-)lines of nbf butterflies each: its
// butterfly j and the LDS base and step of its (sub-)line.  Lanes run along
// the lines when the tile is n-major (INNER), along the butterflies otherwise.
template <int INNER>
__device__ __forceinline__ void locate(const KParams &P, const SubLines &g, bool sub, const FastDiv &fNbf, int nbf,
                                       int e, int &j, int &base, int &step) {
    const int nv = g.nv;
    int v;
    if (INNER) { j = (int)g.fnv.div((unsigned)e); v = e - j * nv; }
    else { v = (int)fNbf.div((unsigned)e); j = e - v * nbf; }
    if (!sub) {
        base = INNER ? v : v * P.m;
        step = INNER ? P.L : 1;
        return;
    }
    int l, c;
    if (INNER) { c = (int)P.fL.div((unsigned)v); l = v - c * P.L; }
    else { l = (int)g.fC.div((unsigned)v); c = v - l * g.C; }
    const int p0 = g.off + c * g.coff;
    base = INNER ? p0 * P.L + l : l * P.m + p0;
    step = INNER ? g.str * P.L : g.str;
}

// Compile-time cos/sin (Taylor series on [-pi, pi], double) for the internal
// twiddles of the composite radices; every use has constant arguments after
// unrolling, so they fold into instruction literals.
constexpr double kPi = 3.14159265358979323846;
constexpr double cx_sin_series(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 24; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_series(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 24; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// exp(SIGN 2 pi i m / R), argument reduced to (-pi, pi]
template <int R, int SIGN>
constexpr double tw_re(int m) {
    m %= R;
    if (2 * m > R) m -= R;
    return cx_cos_series(2.0 * kPi * m / R);
}
template <int R, int SIGN>
constexpr double tw_im(int m) {
    m %= R;
    if (2 * m > R) m -= R;
    return SIGN * cx_sin_series(2.0 * kPi * m / R);
}

// Split of a composite radix R = A * B used by dft (0: a base radix)
constexpr int radix_a(int R) {
    switch (R) {
        case 6: return 2;
        case 8: return 2;
        case 9: return 3;
        case 10: return 2;
        case 12: return 3;
        case 15: return 3;
        default: return 0;
    }
}

// v'[s] = sum_t v[t] exp(SIGN 2 pi i t s / R), in registers.  Radices 2..5 are
// written out; a composite R = A * B runs B-point transforms over t = a + A b,
// the twiddles exp(SIGN 2 pi i a kb / R), and A-point transforms over a, giving
// output kb + B ka.
template <int R, int SIGN>
__device__ __forceinline__ void dft(float2 *v) {
    constexpr float sg = (float)SIGN;
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const float s3 = 0.86602540378443864676f * sg;
        float2 a = v[0], b = v[1], c = v[2];
        float2 t1 = cadd(b, c);
        float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
        float2 d = csub(b, c);
        float2 t3 = make_float2(-s3 * d.y, s3 * d.x);  // i * s3 * (b - c)
        v[0] = cadd(a, t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    } else if constexpr (R == 4) {
        float2 a = v[0], b = v[1], c = v[2], d = v[3];
        float2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
        float2 is3 = make_float2(-sg * s3.y, sg * s3.x);  // SIGN * i * (b - d)
        v[0] = cadd(s0, s2);
        v[1] = cadd(s1, is3);
        v[2] = csub(s0, s2);
        v[3] = csub(s1, is3);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f * sg, s2 = 0.58778525229247312917f * sg;
        float2 a = v[0];
        float2 be = cadd(v[1], v[4]), bm = csub(v[1], v[4]);
        float2 cd = cadd(v[2], v[3]), cm = csub(v[2], v[3]);
        float2 r1 = make_float2(a.x + c1 * be.x + c2 * cd.x, a.y + c1 * be.y + c2 * cd.y);
        float2 r2 = make_float2(a.x + c2 * be.x + c1 * cd.x, a.y + c2 * be.y + c1 * cd.y);
        float2 i1 = make_float2(s1 * bm.x + s2 * cm.x, s1 * bm.y + s2 * cm.y);
        float2 i2 = make_float2(s2 * bm.x - s1 * cm.x, s2 * bm.y - s1 * cm.y);
        v[0] = make_float2(a.x + be.x + cd.x, a.y + be.y + cd.y);
        v[1] = make_float2(r1.x - i1.y, r1.y + i1.x);
        v[4] = make_float2(r1.x + i1.y, r1.y - i1.x);
        v[2] = make_float2(r2.x - i2.y, r2.y + i2.x);
        v[3] = make_float2(r2.x + i2.y, r2.y - i2.x);
    } else {
        constexpr int A = radix_a(R), B = R / A;
        static_assert(A > 1 && A * B == R, "unsupported radix");
        float2 y[R];
#pragma unroll
        for (int a = 0; a < A; ++a) {
            float2 t[B];
#pragma unroll
            for (int b = 0; b < B; ++b) t[b] = v[a + A * b];
            dft<B, SIGN>(t);
#pragma unroll
            for (int kb = 0; kb < B; ++kb) {
                if (a * kb == 0) {
                    y[a * B + kb] = t[kb];
                } else {
                    const float wr = (float)tw_re<R, SIGN>(a * kb), wi = (float)tw_im<R, SIGN>(a * kb);
                    y[a * B + kb] = make_float2(t[kb].x * wr - t[kb].y * wi, t[kb].x * wi + t[kb].y * wr);
                }
            }
        }
#pragma unroll
        for (int kb = 0; kb < B; ++kb) {
            float2 t[A];
#pragma unroll
            for (int a = 0; a < A; ++a) t[a] = y[a * B + kb];
            dft<A, SIGN>(t);
#pragma unroll
            for (int ka = 0; ka < A; ++ka) v[kb + B * ka] = t[ka];
        }
    }
}

// One radix-R Stockham stage over a (sub-)line set of length len.
template <int R, int INNER, int NT, int MAXLEN>
__device__ __forceinline__ void stage(float2 *__restrict__ buf, const KParams &P, const SubLines &g, bool sub, int len,
                                      int Ns, int sign, const float2 *__restrict__ tw, const FastDiv &fNbf,
                                      const FastDiv &fNs) {
    constexpr int MAXB = (MAXLEN / R + NT - 1) / NT;
    const int nbf = len / R;
    const int tot = g.nv * nbf;
    float2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int e = threadIdx.x + b * NT;
        if (e < tot) {
            int j, base, step;
            locate<INNER>(P, g, sub, fNbf, nbf, e, j, base, step);
#pragma unroll
            for (int t = 0; t < R; ++t) v[b][t] = buf[base + (j + t * nbf) * step];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int e = threadIdx.x + b * NT;
        if (e < tot) {
            int j, base, step;
            locate<INNER>(P, g, sub, fNbf, nbf, e, j, base, step);
            const int jq = (int)fNs.div((unsigned)j);
            const int k = j - jq * Ns;
            if (Ns > 1) {
                const float2 *w = tw + k * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) v[b][t] = cmul(v[b][t], w[t - 1]);
            }
            if (sign < 0) dft<R, -1>(v[b]);
            else dft<R, 1>(v[b]);
            const int o = jq * Ns * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) buf[base + (o + t * Ns) * step] = v[b][t];
        }
    }
    __syncthreads();
}

// Q-point DFTs of a prime-factor plan's rows, directly: with s_j = x_j + x_{Q-j},
// d_j = x_j - x_{Q-j} (j = 1..H, H = (Q - 1) / 2),
//   X_k     = A_k + SIGN i B_k,   X_{Q-k} = A_k - SIGN i B_k,
//   A_k = x_0 + sum_j s_j cos(2 pi j k / Q),   B_k = sum_j d_j sin(2 pi j k / Q),
// and X_0 = x_0 + sum_j s_j.  Wave w computes the pairs k = 1 + w G .. w G + G for
// up to 64 rows at once (one row per lane), so k and j are compile-time in its
// unrolled loops and the cosines and sines are instruction literals.  One
// compute-barrier-write round replaces Rader's two (Q - 1)-point FFTs and the
// pointwise step between them (five LDS round trips and their barriers).
template <int Q>
struct DqTable {
    float c[Q], s[Q];
};
template <int Q>
constexpr DqTable<Q> dq_table() {
    DqTable<Q> t{};
    for (int m = 0; m < Q; ++m) {
        t.c[m] = (float)tw_re<Q, 1>(m);
        t.s[m] = (float)tw_im<Q, 1>(m);
    }
    return t;
}

template <int Q, int K0, int G>
__device__ __forceinline__ void dq_acc(const float2 *__restrict__ buf, int base, int step, float2 (&A)[G],
                                       float2 (&B)[G], float2 &X0) {
    constexpr int H = (Q - 1) / 2;
    constexpr DqTable<Q> T = dq_table<Q>();
    const float2 x0 = buf[base];
    X0 = x0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        A[g] = x0;
        B[g] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        const float2 a = buf[base + j * step], b = buf[base + (Q - j) * step];
        const float2 sj = cadd(a, b), dj = csub(a, b);
        if (K0 == 1) X0 = cadd(X0, sj);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int k = K0 + g;
            if (k <= H) {
                const float c = T.c[(j * k) % Q], sn = T.s[(j * k) % Q];
                A[g].x += sj.x * c;
                A[g].y += sj.y * c;
                B[g].x += dj.x * sn;
                B[g].y += dj.y * sn;
            }
        }
    }
}

template <int Q, int K0, int G>
__device__ __forceinline__ void dq_write(float2 *__restrict__ buf, int base, int step, const float2 (&A)[G],
                                         const float2 (&B)[G], float2 X0, float sg) {
    constexpr int H = (Q - 1) / 2;
    if (K0 == 1) buf[base] = X0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int k = K0 + g;
        if (k <= H) {
            buf[base + k * step] = make_float2(A[g].x - sg * B[g].y, A[g].y + sg * B[g].x);
            buf[base + (Q - k) * step] = make_float2(A[g].x + sg * B[g].y, A[g].y - sg * B[g].x);
        }
    }
}

template <int Q, int INNER, int NT>
__device__ __forceinline__ void direct_q_stage(float2 *__restrict__ buf, const KParams &P, int sign) {
    constexpr int H = (Q - 1) / 2;
    constexpr int NW = NT / 64;
    static_assert(NW == 4, "direct DFT stage written for four waves");
    constexpr int G = (H + NW - 1) / NW;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const float sg = (float)sign;
    const int nv = P.L * P.S;
    for (int v0 = 0; v0 < nv; v0 += 64) {
        const int v = v0 + lane;
        const bool act = v < nv;
        int base = 0, step = 1;
        if (act) {
            // row c of tile line l: element t at slot c Q + t
            if (INNER) {
                const int c = (int)P.fL.div((unsigned)v), l = v - c * P.L;
                base = c * Q * P.L + l;
                step = P.L;
            } else {
                const int l = (int)P.fS.div((unsigned)v), c = v - l * P.S;
                base = l * P.m + c * Q;
            }
        }
        float2 A[G], B[G], X0;
        if (act) {
            switch (w) {
                case 0: dq_acc<Q, 1, G>(buf, base, step, A, B, X0); break;
                case 1: dq_acc<Q, 1 + G, G>(buf, base, step, A, B, X0); break;
                case 2: dq_acc<Q, 1 + 2 * G, G>(buf, base, step, A, B, X0); break;
                default: dq_acc<Q, 1 + 3 * G, G>(buf, base, step, A, B, X0); break;
            }
        }
        __syncthreads();   // every wave has read these rows before any overwrites them
        if (act) {
            switch (w) {
                case 0: dq_write<Q, 1, G>(buf, base, step, A, B, X0, sg); break;
                case 1: dq_write<Q, 1 + G, G>(buf, base, step, A, B, X0, sg); break;
                case 2: dq_write<Q, 1 + 2 * G, G>(buf, base, step, A, B, X0, sg); break;
                default: dq_write<Q, 1 + 3 * G, G>(buf, base, step, A, B, X0, sg); break;
            }
        }
    }
    __syncthreads();
}

// Rader's step between the two convolution FFTs.  Row c of line l holds y0 (the
// row's DFT input at index 0) at slot 0 and A = FFT_{Q-1}(a) at slots 1..Q-1,
// a_p = y[g^-p].  Afterwards slot 0 holds X[0] = y0 + sum_p a_p = y0 + A[0] and
// slots 1.. hold A * bhat with y0 added to the zero frequency, so that the
// unnormalised inverse FFT gives X[g^q] = y0 + (a * b)_q at slot 1 + q.
template <int INNER>
__device__ __forceinline__ void rader_pointwise(float2 *buf, const KParams &P) {
    const SubLines &g = P.sub_q;
    const int q1 = P.Q - 1;
    for (int e = threadIdx.x; e < g.nv * q1; e += blockDim.x) {
        int j, base, step;
        locate<INNER>(P, g, true, P.fQ1, q1, e, j, base, step);
        const int id = base + j * step;
        const float2 A = buf[id];
        float2 c = cmul(A, P.bhat[j]);
        if (j == 0) {
            const int id0 = base - step;  // slot 0 of the row
            const float2 y0 = buf[id0];
            buf[id0] = cadd(y0, A);
            c = cadd(c, y0);
        }
        buf[id] = c;
    }
    __syncthreads();
}

__device__ __forceinline__ int remap(int o, int split, int gap) { return o < split ? o : o + gap; }

// position m in the cyclic window [lo, lo + len) of a length-n line (len 0: no window)
__device__ __forceinline__ bool in_window(int m, int lo, int len, int n) {
    if (len <= 0) return true;
    int d = m - lo;
    if (d < 0) d += n;
    return d < len;
}

// The tile of this workgroup, numbered per XCD (xcd_order): consecutive tiles along
// the contiguous index share cache lines in the inner-index passes (four 615-point
// columns are 32 bytes of each 128-byte row segment).  Measured at (359, 71, 359) /
// 615: FFT 12.23 -> 10.60 ms per projector step, CG 3.66 -> 3.38 ms per iteration;
// results unchanged (the tiles are only renumbered).
__device__ __forceinline__ uint3 tile_block() {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned total = gx * gy * gridDim.z;
    const unsigned t = xcd_order(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), total);
    const unsigned q = t / gx;
    return make_uint3(t - q * gx, q % gy, q / gy);
}

// Global coordinates of tile line l; false when the line lies past the set.
__device__ __forceinline__ bool tile_line(const KParams &P, int l, int o2base, int ibase, int bz, long long &o1,
                                          int &o2, int &i) {
    const int lo = (int)P.fLi.div((unsigned)l), li = l - lo * P.Li;
    o2 = o2base + lo;
    i = ibase + li;
    if (o2 >= P.a.c2 || i >= P.a.ci) return false;
    o1 = remap(bz, P.a.o1_split, P.a.o1_gap);
    o2 = remap(o2, P.a.o2_split, P.a.o2_gap);
    return true;
}

// DQ != 0: the instantiation for qdirect plans with prime factor Q = DQ (the S-point
// stages, then direct_q_stage); its own kernel so that the other plans' passes keep
// their register allocation.
template <int INNER, int NT, int MAXLEN, int WAVES, int DQ = 0>
__global__ __launch_bounds__(NT, WAVES) void k_fft_axis(KParams P) {
    extern __shared__ float2 buf[];
    const int kind = DQ ? (int)KIND_RADER : P.kind;
    const int L = P.L, M = P.m, N = P.n;
    const AxisPass &a = P.a;
    const uint3 bk = tile_block();
    const int o2base = INNER ? (int)bk.y * P.Lo : (int)bk.x * P.Lo;
    const int ibase = INNER ? (int)bk.x * P.Li : (int)bk.y;
    const int nv = a.nv < 0 ? N : a.nv;
    const int ns = a.ns < 0 ? N : a.ns;
    // the pair modes' second plane exists for this workgroup's o1 (uniform over the tile)
    const int o1w = remap((int)bk.z, a.o1_split, a.o1_gap);
    const bool has2 = o1w < a.c1_in2, put2 = o1w < a.c1_out2;

    // 64-bit offsets of the tile's lines, once per workgroup (-1: line past the set),
    // and a Rader plan's slot tables (input, then output)
    long long *base_in = reinterpret_cast<long long *>(buf + L * M);
    long long *base_out = base_in + L;
    short *pos_lds = reinterpret_cast<short *>(base_out + L);
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        long long o1;
        int o2, i;
        const bool ok = tile_line(P, l, o2base, ibase, (int)bk.z, o1, o2, i);
        base_in[l] = ok ? o1 * a.is1 + (long long)o2 * a.is2 + (long long)i * a.isi : -1;
        base_out[l] = ok ? o1 * a.os1 + (long long)o2 * a.os2 + (long long)i * a.osi : -1;
    }
    if (kind == KIND_RADER)
        for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) pos_lds[i] = P.pos_in[i];
    __syncthreads();

    // Load: each thread computes the addresses of kLoadBatch elements, issues
    // their global loads, and only then writes them to LDS, so that that many
    // loads per thread are in flight (one at a time, the passes ran at a
    // latency-bound ~2 TB/s).
    const int LN = L * N;
    for (int e0 = threadIdx.x; e0 < LN; e0 += kLoadBatch * NT) {
        long long off[kLoadBatch];
        int line[kLoadBatch], nidx[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int e = e0 + u * NT;
            off[u] = -1;
            line[u] = -1;
            nidx[u] = 0;
            if (e < LN) {
                int l, n;
                if (INNER) { n = (int)P.fL.div((unsigned)e); l = e - n * L; }
                else { l = (int)P.fN.div((unsigned)e); n = e - l * N; }
                int mm = n + a.rot_in;
                if (mm >= N) mm -= N;
                const long long lb = base_in[l];
                if (lb >= 0 && (mm < nv || mm >= N - a.nv2) && in_window(mm, a.ld_lo, a.ld_len, N))
                    off[u] = lb + (long long)mm * P.isn;
                line[u] = l;
                nidx[u] = n;
            }
        }
        float2 val[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) val[u] = make_float2(0.f, 0.f);
        if (a.ld_mode == LD_COMPLEX) {
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (off[u] >= 0) val[u] = static_cast<const float2 *>(a.in)[off[u]];
        } else if (a.ld_mode == LD_REAL) {
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (off[u] >= 0) val[u].x = static_cast<const float *>(a.in)[off[u]];
        } else if (a.ld_mode == LD_REAL_PAIR) {
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (off[u] >= 0) {
                    val[u].x = static_cast<const float *>(a.in)[off[u]];
                    if (has2) val[u].y = static_cast<const float *>(a.in2)[off[u]];
                }
        } else {
            float2 sec[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                sec[u] = make_float2(0.f, 0.f);
                if (off[u] >= 0) {
                    val[u] = static_cast<const float2 *>(a.in)[off[u]];
                    if (has2) sec[u] = static_cast<const float2 *>(a.in2)[off[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                val[u].x -= sec[u].y;
                val[u].y += sec[u].x;
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            if (line[u] >= 0) {
                float2 v = val[u];
                if (kind == KIND_BLUESTEIN) v = cmul(v, P.chirp[nidx[u]]);
                const int p = kind == KIND_RADER ? (int)pos_lds[nidx[u]] : nidx[u];
                buf[lidx<INNER>(line[u], p, L, M)] = v;
            }
        }
    }
    if (kind == KIND_BLUESTEIN) {
        const int pad = M - N;
        for (int e = threadIdx.x; e < L * pad; e += blockDim.x) {
            int l, n;
            if (INNER) { const int q = (int)P.fL.div((unsigned)e); l = e - q * L; n = N + q; }
            else { l = (int)P.fPad.div((unsigned)e); n = N + (e - l * pad); }
            buf[lidx<INNER>(l, n, L, M)] = make_float2(0.f, 0.f);
        }
    }
    __syncthreads();

    for (int s = 0; s < P.nst; ++s) {
        const int gs = P.st_g[s];
        if constexpr (DQ != 0) {
            if (gs == 3) {
                direct_q_stage<DQ, INNER, NT>(buf, P, P.st_sign[s]);
                continue;
            }
        }
        const SubLines &g = gs == 2 ? P.sub_q : gs == 1 ? P.sub_s : P.whole;
        const float2 *tw = P.tw + P.st_tw[s];
#define LAM_STAGE(R)                                                                                           \
    case R:                                                                                                    \
        stage<R, INNER, NT, MAXLEN>(buf, P, g, gs != 0, P.st_len[s], P.st_ns[s], P.st_sign[s], tw, P.fNbf[s], P.fNs[s]); \
        break;
        if (DQ != 0 && P.rad[s] == 15) {
            // qdirect plans with S = 15: the S-point columns in one register stage
            if constexpr (DQ != 0)
                stage<15, INNER, NT, MAXLEN>(buf, P, g, true, P.st_len[s], P.st_ns[s], P.st_sign[s], tw, P.fNbf[s],
                                             P.fNs[s]);
            continue;
        }
        switch (P.rad[s]) {
            LAM_STAGE(2) LAM_STAGE(3) LAM_STAGE(4) LAM_STAGE(5) LAM_STAGE(6) LAM_STAGE(8) LAM_STAGE(9)
            LAM_STAGE(10) LAM_STAGE(12)
            default: break;
        }
#undef LAM_STAGE
        if (s == P.pw_after) {
            if (kind == KIND_RADER) {
                rader_pointwise<INNER>(buf, P);
            } else {
                for (int e = threadIdx.x; e < L * M; e += blockDim.x) {
                    int l, n;
                    if (INNER) { n = (int)P.fL.div((unsigned)e); l = e - n * L; }
                    else { l = (int)P.fM.div((unsigned)e); n = e - l * M; }
                    const int id = lidx<INNER>(l, n, L, M);
                    buf[id] = cmul(buf[id], P.bhat[n]);
                }
                __syncthreads();
            }
        }
    }

    // Store, kLoadBatch elements per thread at a time: their LDS reads are issued
    // together, then their global stores.
    for (int e0 = threadIdx.x; e0 < LN; e0 += kLoadBatch * NT) {
        long long off[kLoadBatch];
        int line[kLoadBatch], nidx[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int e = e0 + u * NT;
            off[u] = -1;
            line[u] = 0;
            nidx[u] = 0;
            if (e < LN) {
                int l, n;
                if (INNER) { n = (int)P.fL.div((unsigned)e); l = e - n * L; }
                else { l = (int)P.fN.div((unsigned)e); n = e - l * N; }
                int mm = n + a.rot_out;
                if (mm >= N) mm -= N;
                const long long lb = base_out[l];
                if (lb >= 0 && (mm < ns || mm >= N - a.ns2) && in_window(mm, a.st_lo, a.st_len, N))
                    off[u] = lb + (long long)mm * P.osn;
                line[u] = l;
                nidx[u] = n;
            }
        }
        float2 val[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int p = kind == KIND_RADER ? (int)pos_lds[N + nidx[u]] : nidx[u];
            val[u] = off[u] >= 0 ? buf[lidx<INNER>(line[u], p, L, M)] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            if (off[u] < 0) continue;
            float2 v = val[u];
            if (kind == KIND_BLUESTEIN) v = cmul(v, P.chirp[nidx[u]]);
            v.x *= a.scale;
            v.y *= a.scale;
            if (a.st_mode == ST_COMPLEX) static_cast<float2 *>(a.out)[off[u]] = v;
            else if (a.st_mode == ST_REAL) static_cast<float *>(a.out)[off[u]] = v.x;
            else if (a.st_mode == ST_IMAG) static_cast<float *>(a.out)[off[u]] = v.y;
            else {
                static_cast<float *>(a.out)[off[u]] = v.x;
                if (put2) static_cast<float *>(a.out2)[off[u]] = v.y;
            }
        }
    }
}

// Radices with a register kernel (dft): the base radices 2..5 and composites
// up to 12.  Fewer, larger stages mean fewer LDS round trips, barriers and
// index splits per element (90 = 10 * 9 in two stages instead of four).
// Measured: radices up to 16 or 20 raise the register count (more waves per
// workgroup stall at the barriers) and gain nothing over 12.
constexpr int kRadices[] = {12, 10, 9, 8, 6, 5, 4, 3, 2};

// fewest stages; among those the largest smallest radix (no trailing radix-2 stage
// where a balanced split exists); radices in non-increasing order
void split_search(int n, int maxr, std::vector<int> &cur, std::vector<int> &best) {
    if (n == 1) {
        const auto smallest = [](const std::vector<int> &v) { return v.empty() ? 0 : v.back(); };
        if (best.empty() || cur.size() < best.size() ||
            (cur.size() == best.size() && smallest(cur) > smallest(best)))
            best = cur;
        return;
    }
    if (!best.empty() && cur.size() + 1 > best.size()) return;
    for (int r : kRadices) {
        if (r > maxr || n % r || (!cur.empty() && r > cur.back())) continue;
        cur.push_back(r);
        split_search(n / r, maxr, cur, best);
        cur.pop_back();
    }
}

// radix list for a 2,3,5-smooth length; false if another prime factor remains
bool factor_len(int n, int *rad, int &nrad) {
    nrad = 0;
    int r = n;
    for (int p : {2, 3, 5}) while (r % p == 0) r /= p;
    if (r != 1) return false;
    if (n == 1) return true;
    std::vector<int> cur, best;
    split_search(n, 12, cur, best);
    if (best.empty() || best.size() > 24) return false;
    for (int x : best) rad[nrad++] = x;
    return true;
}

bool smooth235(int n) {
    for (int p : {2, 3, 5}) while (n % p == 0) n /= p;
    return n == 1;
}

bool is_prime(int q) {
    if (q < 2) return false;
    for (int d = 2; d * d <= q; ++d)
        if (q % d == 0) return false;
    return true;
}

// n = S * Q with S 2,3,5-smooth and Q > 5 prime, Q - 1 2,3,5-smooth (largest such Q)
bool rader_split(int n, int &S, int &Q) {
    for (int q = n; q > 5; --q) {
        if (n % q == 0 && is_prime(q) && smooth235(q - 1) && smooth235(n / q)) {
            S = n / q;
            Q = q;
            return true;
        }
    }
    return false;
}

// Prime factors whose transforms run as a direct DFT stage (direct_q_stage has an
// instantiation for each).  Measured against Rader's convolution at 615 = 15 * 41.
bool direct_q(int q) { return q == 41 || q == 61; }

int primitive_root(int q) {
    for (int g = 2; g < q; ++g) {
        long long x = 1;
        int e = 1;
        for (; e < q - 1; ++e) {
            x = x * g % q;
            if (x == 1) break;
        }
        if (e == q - 1) return g;
    }
    return 1;
}

int inverse_mod(int a, int m) {  // m > 1, gcd(a, m) = 1
    for (int x = 1; x < m; ++x)
        if ((long long)a * x % m == 1) return x;
    return 0;
}

// Prime-factor (Good-Thomas) + Rader tables.  Input element k goes to row
// n1 = k u1 mod S, column n2 = k u2 mod Q (u1 = Q^-1 mod S, u2 = S^-1 mod Q);
// output element k comes from row k mod S, column k mod Q; then
//   X[k] = sum_{n1,n2} x[n1,n2] w_S^{n1 k1} w_Q^{n2 k2}.
// Within a row, column 0 stays at slot 0, input column g^-p goes to slot 1 + p
// and output column g^q comes from slot 1 + q (g a primitive root mod Q).
// A qdirect plan keeps the columns in natural order (slot n1 Q + n2 in, k1 Q + k2 out)
// and has no bhat.
void build_rader(Ctx &cx, FftPlan &plan) {
    const int n = plan.n, S = plan.S, Q = plan.Q, q1 = Q - 1;
    const int g = primitive_root(Q);
    std::vector<int> lg(Q, 0);  // discrete log base g
    {
        long long x = 1;
        for (int e = 0; e < q1; ++e) { lg[(int)x] = e; x = x * g % Q; }
    }
    const int u1 = S > 1 ? inverse_mod(Q % S, S) : 0, u2 = inverse_mod(S % Q, Q);
    std::vector<short> pos(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const int n1 = S > 1 ? (int)((long long)k * u1 % S) : 0;
        const int n2 = (int)((long long)k * u2 % Q);
        const int k1 = k % S, k2 = k % Q;
        if (plan.qdirect) {
            pos[k] = (short)(n1 * Q + n2);
            pos[n + k] = (short)(k1 * Q + k2);
        } else {
            pos[k] = (short)(n1 * Q + (n2 == 0 ? 0 : 1 + (q1 - lg[n2]) % q1));
            pos[n + k] = (short)(k1 * Q + (k2 == 0 ? 0 : 1 + lg[k2]));
        }
    }
    plan.pos.alloc(sizeof(short) * pos.size());
    upload_now(cx, plan.pos.ptr, pos.data(), sizeof(short) * pos.size());
    if (plan.qdirect) return;
    // bhat = FFT_{Q-1}(b) / (Q - 1), b_j = exp(sign 2 pi i g^j / Q), in double
    const long double twopi = 6.283185307179586476925286766559L;
    std::vector<std::complex<double>> b(q1);
    {
        long long x = 1;
        for (int j = 0; j < q1; ++j) {
            long double ang = (long double)plan.sign * twopi * (long double)x / (long double)Q;
            b[j] = std::complex<double>((double)std::cos(ang), (double)std::sin(ang));
            x = x * g % Q;
        }
    }
    std::vector<float2> bh(q1);
    for (int f = 0; f < q1; ++f) {
        std::complex<double> acc = 0.0;
        for (int j = 0; j < q1; ++j) {
            long double ang = -twopi * (long double)((long long)j * f % q1) / (long double)q1;
            acc += b[j] * std::complex<double>((double)std::cos(ang), (double)std::sin(ang));
        }
        acc /= (double)q1;
        bh[f] = make_float2((float)acc.real(), (float)acc.imag());
    }
    plan.bhat.alloc(sizeof(float2) * q1);
    upload_now(cx, plan.bhat.ptr, bh.data(), sizeof(float2) * q1);
}

struct StageDesc {
    int rad, len, sign, g, ns;
};

// The plan's flattened stage list: runs of its radices over a (sub-)line set
// (g: 0 the whole lines, 1 the S-point columns, 2 the (Q - 1)-point rows of a
// Rader plan, 3 the direct Q-point DFT of a qdirect plan's rows); pw_after is the
// stage after which Bluestein's or Rader's pointwise product runs (-1: none).
std::vector<StageDesc> stage_list(const FftPlan &plan, int &pw_after) {
    std::vector<StageDesc> st;
    pw_after = -1;
    auto run = [&](int first, int count, int len, int sign, int g) {
        int Ns = 1;
        for (int s = first; s < first + count; ++s) {
            st.push_back({plan.rad[s], len, sign, g, Ns});
            Ns *= plan.rad[s];
        }
    };
    if (plan.qdirect) {
        run(0, plan.nrad_s, plan.S, plan.sign, 1);
        st.push_back({plan.Q, plan.Q, plan.sign, 3, 1});
    } else if (plan.rader) {
        const int nq = plan.nrad - plan.nrad_s;
        run(0, plan.nrad_s, plan.S, plan.sign, 1);
        run(plan.nrad_s, nq, plan.Q - 1, -1, 2);
        pw_after = (int)st.size() - 1;
        run(plan.nrad_s, nq, plan.Q - 1, +1, 2);
    } else if (plan.blue) {
        run(0, plan.nrad, plan.m, -1, 0);
        pw_after = (int)st.size() - 1;
        run(0, plan.nrad, plan.m, +1, 0);
    } else {
        run(0, plan.nrad, plan.m, plan.sign, 0);
    }
    LAM_REQUIRE(st.size() <= 24, LAM_ERR_ARG, "FFT plan has too many stages");
    return st;
}

// Twiddles of every stage, exp(sign 2 pi i t k / (Ns R)) for k < Ns and
// t = 1..R-1, laid out [k][t - 1] so that a butterfly reads its R - 1 values
// from one base address (computed in long double, rounded once).
void build_twiddles(Ctx &cx, FftPlan &plan) {
    int pw;
    const auto st = stage_list(plan, pw);
    const long double twopi = 6.283185307179586476925286766559L;
    std::vector<float2> tw;
    for (size_t i = 0; i < st.size(); ++i) {
        plan.tw_off[i] = (int)tw.size();
        const StageDesc &d = st[i];
        if (d.ns <= 1) continue;
        const long long span = (long long)d.ns * d.rad;
        for (int k = 0; k < d.ns; ++k)
            for (int t = 1; t < d.rad; ++t) {
                const long double ang = (long double)d.sign * twopi * (long double)((long long)t * k % span) /
                                        (long double)span;
                tw.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
            }
    }
    if (tw.empty()) tw.push_back(make_float2(1.f, 0.f));
    plan.tw.alloc(sizeof(float2) * tw.size());
    upload_now(cx, plan.tw.ptr, tw.data(), sizeof(float2) * tw.size());
}

}  // namespace

FftPlan &get_fft_plan(Ctx &cx, int n, int sign) {
    auto key = std::make_pair(n, sign);
    auto it = cx.fft_plans.find(key);
    if (it != cx.fft_plans.end()) return *it->second;
    LAM_REQUIRE(n >= 1, LAM_ERR_ARG, "FFT length must be positive");
    auto plan = std::make_unique<FftPlan>();
    plan->n = n;
    plan->sign = sign;
    int rad[24], nrad;
    int S = 1, Q = 1;
    if (n == 1) {
        plan->m = 1;
        plan->nrad = 0;
    } else if (factor_len(n, rad, nrad)) {
        plan->m = n;
        plan->nrad = nrad;
        for (int i = 0; i < nrad; ++i) plan->rad[i] = rad[i];
    } else if (n <= kMaxLen && rader_split(n, S, Q)) {
        plan->rader = true;
        plan->qdirect = direct_q(Q) && n <= kLdsComplex;
        plan->m = n;
        plan->S = S;
        plan->Q = Q;
        int rs[24], ns = 0, rq[24], nq = 0;
        if (plan->qdirect && S == 15) {
            // one register stage (the qdirect passes have a radix-15 stage; measured at 615:
            // FFT 8.90 -> 8.69 ms per projector step against radices 5 and 3)
            rs[0] = 15;
            ns = 1;
        } else if (S > 1) {
            factor_len(S, rs, ns);
        }
        if (!plan->qdirect) factor_len(Q - 1, rq, nq);
        plan->nrad_s = ns;
        plan->nrad = ns + nq;
        for (int i = 0; i < ns; ++i) plan->rad[i] = rs[i];
        for (int i = 0; i < nq; ++i) plan->rad[ns + i] = rq[i];
        build_rader(cx, *plan);
        build_twiddles(cx, *plan);
        FftPlan &ref = *plan;
        cx.fft_plans.emplace(key, std::move(plan));
        return ref;
    } else {
        plan->blue = true;
        int m = 2 * n - 1;
        while (!smooth235(m)) ++m;
        plan->m = m;
        factor_len(m, rad, nrad);
        plan->nrad = nrad;
        for (int i = 0; i < nrad; ++i) plan->rad[i] = rad[i];
    }
    const int m = plan->m;
    LAM_REQUIRE(m <= kMaxLen, LAM_ERR_ARG, "FFT length " + std::to_string(n) + " exceeds the LDS transform limit");
    const long double twopi = 6.283185307179586476925286766559L;
    if (plan->blue) {
        std::vector<std::complex<double>> ch(n), b(m, 0.0), bh(m);
        std::vector<float2> chf(n), bhf(m);
        for (int k = 0; k < n; ++k) {
            unsigned long long q = (unsigned long long)k * (unsigned long long)k % (2ull * (unsigned long long)n);
            long double ang = (long double)sign * (twopi / 2) * (long double)q / (long double)n;
            ch[k] = std::complex<double>((double)std::cos(ang), (double)std::sin(ang));
            chf[k] = make_float2((float)ch[k].real(), (float)ch[k].imag());
        }
        for (int k = 0; k < n; ++k) {
            b[k] = std::conj(ch[k]);
            if (k) b[m - k] = std::conj(ch[k]);
        }
        // direct DFT in double with an exact root table (m <= 4096)
        std::vector<std::complex<double>> rt(m);
        for (int k = 0; k < m; ++k) {
            long double ang = -twopi * (long double)k / (long double)m;
            rt[k] = std::complex<double>((double)std::cos(ang), (double)std::sin(ang));
        }
        for (int f = 0; f < m; ++f) {
            std::complex<double> acc = 0.0;
            long long idx = 0;
            for (int k = 0; k < m; ++k) {
                if (b[k] != 0.0) acc += b[k] * rt[idx];
                idx += f;
                if (idx >= m) idx -= m;
            }
            bh[f] = acc / (double)m;
            bhf[f] = make_float2((float)bh[f].real(), (float)bh[f].imag());
        }
        plan->chirp.alloc(sizeof(float2) * n);
        plan->bhat.alloc(sizeof(float2) * m);
        upload_now(cx, plan->chirp.ptr, chf.data(), sizeof(float2) * n);
        upload_now(cx, plan->bhat.ptr, bhf.data(), sizeof(float2) * m);
    }
    build_twiddles(cx, *plan);
    FftPlan &ref = *plan;
    cx.fft_plans.emplace(key, std::move(plan));
    return ref;
}

void run_axis(Ctx &cx, const FftPlan &plan, AxisPass a) {
    if (a.c1 <= 0 || a.c2 <= 0 || a.ci <= 0) return;
    KParams P;
    P.a = a;
    P.n = plan.n;
    P.m = plan.m;
    P.kind = plan.rader ? KIND_RADER : plan.blue ? KIND_BLUESTEIN : KIND_STOCKHAM;
    P.sign = plan.sign;
    P.chirp = plan.chirp.as<float2>();
    P.bhat = plan.bhat.as<float2>();
    const bool inner = a.ci > 1;
    const int cap = std::max(1, std::min(kLdsComplex / plan.m, kMaxLines));
    dim3 grid;
    if (inner) {
        P.Li = std::min({a.ci, cap, kInnerLines});
        P.Lo = std::max(1, std::min(a.c2, cap / P.Li));
        grid = dim3((unsigned)((a.ci + P.Li - 1) / P.Li), (unsigned)((a.c2 + P.Lo - 1) / P.Lo), (unsigned)a.c1);
    } else {
        P.Li = 1;
        P.Lo = std::min({a.c2, cap, 32});
        grid = dim3((unsigned)((a.c2 + P.Lo - 1) / P.Lo), (unsigned)a.ci, (unsigned)a.c1);
    }
    P.L = P.Li * P.Lo;
    LAM_REQUIRE(a.isn > -(1ll << 31) && a.isn < (1ll << 31) && a.osn > -(1ll << 31) && a.osn < (1ll << 31),
                LAM_ERR_ARG, "FFT element stride exceeds 32 bits");
    P.isn = (int)a.isn;
    P.osn = (int)a.osn;
    P.fL = FastDiv((unsigned)P.L);
    P.fN = FastDiv((unsigned)plan.n);
    P.fM = FastDiv((unsigned)plan.m);
    P.fLi = FastDiv((unsigned)P.Li);
    P.fPad = FastDiv((unsigned)std::max(1, plan.m - plan.n));
    // the plan's flattened stage list with its twiddle tables
    int pw_after;
    const auto st = stage_list(plan, pw_after);
    P.nst = (int)st.size();
    P.pw_after = pw_after;
    P.tw = plan.tw.as<float2>();
    for (int i = 0; i < P.nst; ++i) {
        P.rad[i] = st[i].rad;
        P.st_len[i] = st[i].len;
        P.st_sign[i] = st[i].sign;
        P.st_g[i] = st[i].g;
        P.st_ns[i] = st[i].ns;
        P.st_tw[i] = plan.tw_off[i];
        P.fNbf[i] = FastDiv((unsigned)(st[i].len / st[i].rad));
        P.fNs[i] = FastDiv((unsigned)st[i].ns);
    }
    P.whole.nv = P.L; P.whole.C = 1; P.whole.off = 0; P.whole.coff = 0; P.whole.str = 1;
    P.whole.fnv = FastDiv((unsigned)P.L); P.whole.fC = FastDiv(1u);
    P.S = P.Q = 1;
    P.pos_in = P.pos_out = nullptr;
    if (plan.rader) {
        P.S = plan.S;
        P.Q = plan.Q;
        P.pos_in = plan.pos.as<short>();
        P.pos_out = P.pos_in + plan.n;
        P.fQ1 = FastDiv((unsigned)(plan.Q - 1));
        P.fS = FastDiv((unsigned)plan.S);
        // S-point transforms: Q columns per line, element t of column c at t * Q + c
        P.sub_s.nv = P.L * plan.Q; P.sub_s.C = plan.Q; P.sub_s.off = 0; P.sub_s.coff = 1; P.sub_s.str = plan.Q;
        P.sub_s.fnv = FastDiv((unsigned)P.sub_s.nv); P.sub_s.fC = FastDiv((unsigned)plan.Q);
        // Q - 1 point convolutions: S rows per line, element t of row c at c * Q + 1 + t
        P.sub_q.nv = P.L * plan.S; P.sub_q.C = plan.S; P.sub_q.off = 1; P.sub_q.coff = plan.Q; P.sub_q.str = 1;
        P.sub_q.fnv = FastDiv((unsigned)P.sub_q.nv); P.sub_q.fC = FastDiv((unsigned)plan.S);
    }
    LAM_REQUIRE(grid.y <= 65535 && grid.z <= 65535, LAM_ERR_ARG, "FFT pass grid exceeds launch limits");
    const size_t lds = (size_t)P.L * plan.m * sizeof(float2) + 2 * sizeof(long long) * (size_t)P.L +
                       (plan.rader ? 2 * sizeof(short) * (size_t)plan.n : 0);
    TimedScope ts(cx, "fft_axis");
    if (plan.qdirect) {
        LAM_REQUIRE(plan.m <= kLdsComplex && (plan.Q == 41 || plan.Q == 61), LAM_ERR_STATE, "no direct DFT pass for this plan");
        if (plan.Q == 41) {
            if (inner)
                hipLaunchKernelGGL((k_fft_axis<1, kThreads, kLdsComplex, kFftWaves, 41>), grid, dim3(kThreads), lds, cx.stream, P);
            else
                hipLaunchKernelGGL((k_fft_axis<0, kThreads, kLdsComplex, kFftWaves, 41>), grid, dim3(kThreads), lds, cx.stream, P);
        } else {
            if (inner)
                hipLaunchKernelGGL((k_fft_axis<1, kThreads, kLdsComplex, kFftWaves, 61>), grid, dim3(kThreads), lds, cx.stream, P);
            else
                hipLaunchKernelGGL((k_fft_axis<0, kThreads, kLdsComplex, kFftWaves, 61>), grid, dim3(kThreads), lds, cx.stream, P);
        }
    } else if (plan.m <= kLdsComplex) {
        if (inner)
            hipLaunchKernelGGL((k_fft_axis<1, kThreads, kLdsComplex, kFftWaves>), grid, dim3(kThreads), lds, cx.stream, P);
        else
            hipLaunchKernelGGL((k_fft_axis<0, kThreads, kLdsComplex, kFftWaves>), grid, dim3(kThreads), lds, cx.stream, P);
    } else {
        if (inner)
            hipLaunchKernelGGL((k_fft_axis<1, kThreads, kMaxLen, 1>), grid, dim3(kThreads), lds, cx.stream, P);
        else
            hipLaunchKernelGGL((k_fft_axis<0, kThreads, kMaxLen, 1>), grid, dim3(kThreads), lds, cx.stream, P);
    }
    check_launch();
}

// ------------------------------------------------------------------ Toeplitz spectral step
// The stages of a 2,3,5-smooth plan, for kernels that run a whole transform on
// LDS-resident lines between other work.
struct StageList {
    int nst = 0;
    int rad[24] = {}, len[24] = {}, ns[24] = {}, sign[24] = {}, tw[24] = {};
    const float2 *twp = nullptr;
    FastDiv fNbf[24], fNs[24];
};

namespace {

StageList make_stage_list(const FftPlan &plan) {
    LAM_REQUIRE(!plan.blue && !plan.rader, LAM_ERR_STATE, "fused transform needs a 2,3,5-smooth length");
    int pw;
    const auto st = stage_list(plan, pw);
    StageList S;
    S.nst = (int)st.size();
    S.twp = plan.tw.as<float2>();
    for (int i = 0; i < S.nst; ++i) {
        S.rad[i] = st[i].rad;
        S.len[i] = st[i].len;
        S.ns[i] = st[i].ns;
        S.sign[i] = st[i].sign;
        S.tw[i] = plan.tw_off[i];
        S.fNbf[i] = FastDiv((unsigned)(st[i].len / st[i].rad));
        S.fNs[i] = FastDiv((unsigned)st[i].ns);
    }
    return S;
}

template <int NT, int MAXLEN>
__device__ __forceinline__ void run_stages(float2 *buf, const KParams &P, const SubLines &g, const StageList &S) {
    for (int s = 0; s < S.nst; ++s) {
        const float2 *tw = S.twp + S.tw[s];
#define LAM_FSTAGE(R)                                                                                            \
    case R:                                                                                                      \
        stage<R, 0, NT, MAXLEN>(buf, P, g, false, S.len[s], S.ns[s], S.sign[s], tw, S.fNbf[s], S.fNs[s]);       \
        break;
        switch (S.rad[s]) {
            LAM_FSTAGE(2) LAM_FSTAGE(3) LAM_FSTAGE(4) LAM_FSTAGE(5) LAM_FSTAGE(6) LAM_FSTAGE(8) LAM_FSTAGE(9)
            LAM_FSTAGE(10) LAM_FSTAGE(12)
            default: break;
        }
#undef LAM_FSTAGE
    }
}

// X0, X1 from the packed spectrum Z = FFT(x0 + i x1) at f (a) and -f (b).
__device__ __forceinline__ void unpack_pair(float2 a, float2 b, float2 &X0, float2 &X1) {
    X0 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    const float2 t = make_float2(a.x - b.x, a.y + b.y);
    X1 = make_float2(0.5f * t.y, -0.5f * t.x);
}

// Y = K X with the real symmetric 3 x 3 spectrum (K00,K01),(K02,K11),(K12,K22), packed
// back as (Y0 + i Y1, Y2) for the inverse transform's real-pair output.
__device__ __forceinline__ void toep_block(float2 k0, float2 k1, float2 k2, float2 X0, float2 X1, float2 X2,
                                           float2 &oa, float2 &ob) {
    float2 Y0, Y1, Y2;
    Y0.x = k0.x * X0.x + k0.y * X1.x + k1.x * X2.x;
    Y0.y = k0.x * X0.y + k0.y * X1.y + k1.x * X2.y;
    Y1.x = k0.y * X0.x + k1.y * X1.x + k2.x * X2.x;
    Y1.y = k0.y * X0.y + k1.y * X1.y + k2.x * X2.y;
    Y2.x = k1.x * X0.x + k2.x * X1.x + k2.y * X2.x;
    Y2.y = k1.x * X0.y + k2.x * X1.y + k2.y * X2.y;
    oa = make_float2(Y0.x - Y1.y, Y0.y + Y1.x);
    ob = Y2;
}

// The normal operator's spectral step along the contiguous axis, fused: for the
// spectrum row r = (z, y) (already transformed along z and y) and its mirror row
// (-z, -y), the workgroup loads the four lines za[r], zb[r], za[mr], zb[mr]
// (positions x < nvx; the rest of each line is the zero padding), transforms them
// along x, multiplies by the kernel spectrum, transforms back and stores x < nvx.
// The unpacking of x0 + i x1 needs Za at -f, which lies on the mirror row: the two
// rows are done together.  The kernel spectra are real and even, K(-f) = K(f), so
// only row r's kernel row is read.  Rows past their mirror are skipped.
template <int NT, int MAXLEN>
__global__ __launch_bounds__(NT) void k_toep_xfused(StageList F, StageList I, const float2 *__restrict__ za,
                                                    const float2 *__restrict__ zb, const float2 *__restrict__ K,
                                                    float2 *__restrict__ wa, float2 *__restrict__ wb, int L1, int L2,
                                                    int L3, int nvx, FastDiv fL3) {
    extern __shared__ float2 buf[];
    const int r = (int)blockIdx.x;
    const int z = r / L2, y = r - z * L2;
    const int mr = (z ? L1 - z : 0) * L2 + (y ? L2 - y : 0);
    if (mr < r) return;
    const bool self = mr == r;
    const int nl = self ? 2 : 4;
    const long long rb = (long long)r * L3, mb = (long long)mr * L3;
    for (int e = threadIdx.x; e < nl * L3; e += NT) {
        const int l = (int)fL3.div((unsigned)e), n = e - l * L3;
        const float2 *src = (l & 1) ? zb : za;
        buf[e] = n < nvx ? src[(l < 2 ? rb : mb) + n] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    KParams P;
    P.m = L3;
    P.L = nl;
    SubLines g;
    g.nv = nl;
    run_stages<NT, MAXLEN>(buf, P, g, F);
    constexpr int kMaxX = (MAXLEN / 4 + NT - 1) / NT;   // x positions per thread (4 lines of L3 <= MAXLEN / 4)
    float2 oa[kMaxX], ob[kMaxX], oam[kMaxX], obm[kMaxX];
#pragma unroll
    for (int u = 0; u < kMaxX; ++u) {
        const int x = threadIdx.x + u * NT;
        if (x < L3) {
            const int xm = x ? L3 - x : 0;
            const float2 *kr = K + 3 * (rb + x);
            float2 X0, X1;
            if (self) {
                unpack_pair(buf[x], buf[xm], X0, X1);
                toep_block(kr[0], kr[1], kr[2], X0, X1, buf[L3 + x], oa[u], ob[u]);
            } else {
                unpack_pair(buf[x], buf[2 * L3 + xm], X0, X1);
                toep_block(kr[0], kr[1], kr[2], X0, X1, buf[L3 + x], oa[u], ob[u]);
                const float2 *km = K + 3 * (rb + xm);   // K(mr, x) = K(r, -x)
                unpack_pair(buf[2 * L3 + x], buf[xm], X0, X1);
                toep_block(km[0], km[1], km[2], X0, X1, buf[3 * L3 + x], oam[u], obm[u]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMaxX; ++u) {
        const int x = threadIdx.x + u * NT;
        if (x < L3) {
            buf[x] = oa[u];
            buf[L3 + x] = ob[u];
            if (!self) {
                buf[2 * L3 + x] = oam[u];
                buf[3 * L3 + x] = obm[u];
            }
        }
    }
    __syncthreads();
    run_stages<NT, MAXLEN>(buf, P, g, I);
    for (int e = threadIdx.x; e < nl * L3; e += NT) {
        const int l = (int)fL3.div((unsigned)e), n = e - l * L3;
        if (n < nvx) ((l & 1) ? wb : wa)[(l < 2 ? rb : mb) + n] = buf[e];
    }
}

}  // namespace

void toep_spectral_x(Ctx &cx, const float2 *za, const float2 *zb, const float2 *K, float2 *wa, float2 *wb, int L1,
                     int L2, int L3, int nvx) {
    const FftPlan &pf = get_fft_plan(cx, L3, -1), &pi = get_fft_plan(cx, L3, +1);
    const StageList F = make_stage_list(pf), I = make_stage_list(pi);
    const long long rows = (long long)L1 * L2;
    LAM_REQUIRE(rows < (1ll << 31), LAM_ERR_ARG, "Toeplitz grid too large");
    const size_t lds = sizeof(float2) * 4 * (size_t)L3;
    TimedScope ts(cx, "toep_xfused");
    if (4 * L3 <= kLdsComplex) {
        hipLaunchKernelGGL((k_toep_xfused<kThreads, kLdsComplex>), dim3((unsigned)rows), dim3(kThreads), lds, cx.stream,
                           F, I, za, zb, K, wa, wb, L1, L2, L3, nvx, FastDiv((unsigned)L3));
    } else {
        LAM_REQUIRE(4 * L3 <= kFusedMaxLen, LAM_ERR_ARG, "Toeplitz grid line too long for the fused step");
        const void *fn = (const void *)&k_toep_xfused<kThreads, kFusedMaxLen>;
        if (lds > 64 * 1024) LAM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_toep_xfused<kThreads, kFusedMaxLen>), dim3((unsigned)rows), dim3(kThreads), lds,
                           cx.stream, F, I, za, zb, K, wa, wb, L1, L2, L3, nvx, FastDiv((unsigned)L3));
    }
    check_launch();
}

void det_fft2(Ctx &cx, const void *in, bool in_real, void *out, bool out_real, int P, int R, int C, int sign,
              float scale, float2 *tmp) {
    // pass 1: along columns (contiguous), lines (p, r)
    {
        AxisPass a;
        a.in = in;
        a.ld_mode = in_real ? LD_REAL : LD_COMPLEX;
        a.is1 = (long long)R * C; a.is2 = C; a.isi = 0; a.isn = 1;
        a.out = tmp;
        a.st_mode = ST_COMPLEX;
        a.os1 = (long long)R * C; a.os2 = C; a.osi = 0; a.osn = 1;
        a.c1 = P; a.c2 = R; a.ci = 1;
        a.rot_in = C / 2;
        a.rot_out = C / 2;
        run_axis(cx, get_fft_plan(cx, C, sign), a);
    }
    // pass 2: along rows, lines (p, c), tiled over c
    {
        AxisPass a;
        a.in = tmp;
        a.ld_mode = LD_COMPLEX;
        a.is1 = (long long)R * C; a.is2 = 0; a.isi = 1; a.isn = C;
        a.out = out;
        a.st_mode = out_real ? ST_REAL : ST_COMPLEX;
        a.os1 = (long long)R * C; a.os2 = 0; a.osi = 1; a.osn = C;
        a.c1 = P; a.c2 = 1; a.ci = C;
        a.rot_in = R / 2;
        a.rot_out = R / 2;
        a.scale = scale;
        run_axis(cx, get_fft_plan(cx, R, sign), a);
    }
}

namespace {

void det_pairs(Ctx &cx, const float *proj_in, float *proj_out, float2 *spec, int P, int R, int C, int sign,
               float scale) {
    LAM_REQUIRE((R & 1) && (C & 1), LAM_ERR_ARG, "paired detector transforms need odd plane sizes");
    const bool to_real = proj_out != nullptr;
    const long long RC = (long long)R * C;
    const int pairs = (P + 1) / 2, full = P / 2;   // pairs with a second plane: o1 < full
    // pass 1: along columns (contiguous), lines (pair, r); pair m's spectrum lives in plane 2m
    {
        AxisPass a;
        if (to_real) {
            a.in = spec; a.in2 = spec + RC; a.ld_mode = LD_COMPLEX_PAIR;
        } else {
            a.in = proj_in; a.in2 = proj_in + RC; a.ld_mode = LD_REAL_PAIR;
        }
        a.c1_in2 = full;
        a.is1 = 2 * RC; a.is2 = C; a.isi = 0; a.isn = 1;
        a.out = spec;
        a.st_mode = ST_COMPLEX;
        a.os1 = 2 * RC; a.os2 = C; a.osi = 0; a.osn = 1;
        a.c1 = pairs; a.c2 = R; a.ci = 1;
        a.rot_in = C / 2;
        a.rot_out = C / 2;
        run_axis(cx, get_fft_plan(cx, C, sign), a);
    }
    // pass 2: along rows, lines (pair, c), tiled over c
    {
        AxisPass a;
        a.in = spec;
        a.ld_mode = LD_COMPLEX;
        a.is1 = 2 * RC; a.is2 = 0; a.isi = 1; a.isn = C;
        if (to_real) {
            a.out = proj_out; a.out2 = proj_out + RC; a.st_mode = ST_REAL_PAIR;
            a.c1_out2 = full;
        } else {
            a.out = spec; a.st_mode = ST_COMPLEX;
        }
        a.os1 = 2 * RC; a.os2 = 0; a.osi = 1; a.osn = C;
        a.c1 = pairs; a.c2 = 1; a.ci = C;
        a.rot_in = R / 2;
        a.rot_out = R / 2;
        a.scale = scale;
        run_axis(cx, get_fft_plan(cx, R, sign), a);
    }
}

}  // namespace

void det_fft2_real_pairs(Ctx &cx, const float *proj, float2 *spec, int P, int R, int C, int sign) {
    det_pairs(cx, proj, nullptr, spec, P, R, C, sign, 1.0f);
}

void det_ifft2_real_pairs(Ctx &cx, float2 *spec, float *proj, int P, int R, int C, int sign, float scale) {
    det_pairs(cx, nullptr, proj, spec, P, R, C, sign, scale);
}

void fft3d(Ctx &cx, float2 *data, int A, int B, int Cn, int nc, int sign) {
    const long long sC = nc, sB = (long long)Cn * nc, sA = (long long)B * Cn * nc;
    {   // axis C (x): lines (a, b) x components
        AxisPass p;
        p.in = data; p.out = data;
        p.is1 = p.os1 = sA; p.is2 = p.os2 = sB; p.isi = p.osi = 1; p.isn = p.osn = sC;
        p.c1 = A; p.c2 = B; p.ci = nc;
        run_axis(cx, get_fft_plan(cx, Cn, sign), p);
    }
    {   // axis B (y): lines (a, c*nc)
        AxisPass p;
        p.in = data; p.out = data;
        p.is1 = p.os1 = sA; p.is2 = p.os2 = 0; p.isi = p.osi = 1; p.isn = p.osn = sB;
        p.c1 = A; p.c2 = 1; p.ci = Cn * nc;
        run_axis(cx, get_fft_plan(cx, B, sign), p);
    }
    {   // axis A (z): lines (b, c*nc)
        AxisPass p;
        p.in = data; p.out = data;
        p.is1 = p.os1 = 0; p.is2 = p.os2 = sB; p.isi = p.osi = 1; p.isn = p.osn = sA;
        p.c1 = 1; p.c2 = B; p.ci = Cn * nc;
        run_axis(cx, get_fft_plan(cx, A, sign), p);
    }
}

}  // namespace lam
