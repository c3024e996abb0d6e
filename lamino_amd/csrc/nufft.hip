// Polar-grid NUFFT on gfx950: deapodise + zero-pad, fine-grid FFT,
// exponential-of-semicircle interpolation / spreading with the point
// coordinates generated analytically per (projection, row, col).
//
// For real volumes the plane data are Hermitian: the point (R-1-j, C-1-k) is
// the negative of (j, k), so c(-q) = conj(c(q)).  The projector kernels
// therefore visit only the first half of every plane:
//   - interpolation writes c(q) and conj(c(q)) to the mirrored slot;
//   - spreading of any plane data followed by Re() equals spreading
//     c(q) + conj(c(-q)) over the half set (exact, no symmetry assumed).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>

#include "fastdiv.h"
#include "nufft.h"


namespace lam {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ int wrapi(int i, int n) {
    if (i < 0) i += n;
    else if (i >= n) i -= n;
    return i;
}

// Coordinate arithmetic is kept free of FMA contraction so that every kernel
// that places a point (binning, spreading, interpolation) computes bitwise the
// same fine coordinate and first kernel cell.

// Folds t into one period and returns the first kernel cell ceil(t - W/2).
// The fold multiplies by the reciprocal 1/nf: at an exact multiple of nf it
// may land on -0 ulp or nf, which every caller wraps (wrap0) identically.
__device__ __forceinline__ int first_cell(double &t, int nf, double rnf, int W) {
#pragma clang fp contract(off)
    const double fl = floor(t * rnf);
    t -= fl * (double)nf;
    return (int)ceil(t - 0.5 * W);
}

// ES kernel weights for one axis: cells i0 .. i0+W-1 around the folded coordinate t.
template <int W>
__device__ __forceinline__ void es_weights(double t, int i0, float beta, float *wts) {
    const float z0 = (float)(((double)i0 - t) * (2.0 / W));
    const float dz = 2.0f / (float)W;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const float z = z0 + dz * (float)m;
        const float r = 1.0f - z * z;
        // r in (0, 1]: the hardware square root (no denormal rescaling needed)
        wts[m] = r > 0.f ? __expf(beta * (__builtin_amdgcn_sqrtf(r) - 1.0f)) : 0.f;
    }
}

struct PointFrame {
    double dX, dY, k0, j0, sx, sy, sz;  // qX step, qY step, centres, fine scales nf/(2 pi)
    double rx, ry, rz;                  // 1 / nf per axis
};

__device__ __forceinline__ void point_coords(const Geo &g, const PointFrame &f, int j, int k, double &tx, double &ty,
                                             double &tz) {
#pragma clang fp contract(off)
    const double qX = ((double)k - f.k0) * f.dX;
    const double qY = ((double)j - f.j0) * f.dY;
    tx = (qX * g.ux + qY * g.vx) * f.sx;
    ty = (qY * g.vy) * f.sy;
    tz = (qX * g.uz + qY * g.vz) * f.sz;
}

// Placement of one point: folded coordinates, first cells, weights.
template <int W>
struct Placed {
    int ix, iy, iz;
    float wx[W], wy[W], wz[W];
};

template <int W>
__device__ __forceinline__ void place(const Geo &g, const PointFrame &fr, int j, int k, int nfz, int nfy, int nfx,
                                      float beta, Placed<W> &pt) {
    double tx, ty, tz;
    point_coords(g, fr, j, k, tx, ty, tz);
    pt.ix = first_cell(tx, nfx, fr.rx, W);
    pt.iy = first_cell(ty, nfy, fr.ry, W);
    pt.iz = first_cell(tz, nfz, fr.rz, W);
    es_weights<W>(tx, pt.ix, beta, pt.wx);
    es_weights<W>(ty, pt.iy, beta, pt.wy);
    es_weights<W>(tz, pt.iz, beta, pt.wz);
}

// Tiling of the fine grid for the spreader: tile (tz,ty,tx) owns the points
// whose first kernel cell (wrapped into [0,nf)) lies in it, and accumulates
// them over a (T+W-1)^3 halo box in LDS.
struct BinParams {
    int T[3];    // tile extent z, y, x
    int nt[3];   // tiles per axis
    FastDiv fhx, fhyx;   // halo box row length HX and plane size HY * HX
};

__device__ __forceinline__ int wrap0(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ------------------------------------------------------------------ deapodise + pad
// fine[g][c] = vol_c[i] * corr, array index i <-> mode i - n/2 <-> fine index
// mode mod nf.  Only the mode region is written: the pruned fine-grid FFT
// never reads the rest.
template <int NC, bool REAL>
__global__ __launch_bounds__(256) void k_deapod_pad(float2 *__restrict__ fine, const float *__restrict__ r0,
                                                    const float *__restrict__ r1, const float *__restrict__ r2,
                                                    const float2 *__restrict__ cin, int n1, int n2, int n3,
                                                    int nfz, int nfy, int nfx, const float *__restrict__ cz,
                                                    const float *__restrict__ cy, const float *__restrict__ cxr) {
    const long long total = (long long)n1 * n2 * n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(id % n3);
        const int iy = (int)((id / n3) % n2);
        const int iz = (int)(id / ((long long)n3 * n2));
        int gz = iz - n1 / 2; if (gz < 0) gz += nfz;
        int gy = iy - n2 / 2; if (gy < 0) gy += nfy;
        int gx = ix - n3 / 2; if (gx < 0) gx += nfx;
        const long long fi = ((long long)gz * nfy + gy) * nfx + gx;
        const float cf = cz[iz] * cy[iy] * cxr[ix];
        if constexpr (REAL) {
            fine[fi * NC] = make_float2(r0[id] * cf, 0.f);
            if constexpr (NC == 3) {
                fine[fi * NC + 1] = make_float2(r1[id] * cf, 0.f);
                fine[fi * NC + 2] = make_float2(r2[id] * cf, 0.f);
            }
        } else {
            fine[fi * NC] = cscale(cin[id], cf);
        }
    }
}

// ------------------------------------------------------------------ crop + deapodise
// REAL: out_c[i] (+)= Re(fine[g][c] * corr) * scale; else complex out.
template <int NC, bool REAL>
__global__ __launch_bounds__(256) void k_crop_deapod(const float2 *__restrict__ fine, float *__restrict__ o0,
                                                     float *__restrict__ o1, float *__restrict__ o2,
                                                     float2 *__restrict__ cout, int n1, int n2, int n3, int nfz,
                                                     int nfy, int nfx, const float *__restrict__ cz,
                                                     const float *__restrict__ cy, const float *__restrict__ cxr,
                                                     float scale, int accumulate) {
    const long long total = (long long)n1 * n2 * n3;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(id % n3);
        const int iy = (int)((id / n3) % n2);
        const int iz = (int)(id / ((long long)n3 * n2));
        int gz = iz - n1 / 2; if (gz < 0) gz += nfz;
        int gy = iy - n2 / 2; if (gy < 0) gy += nfy;
        int gx = ix - n3 / 2; if (gx < 0) gx += nfx;
        const long long fi = ((long long)gz * nfy + gy) * nfx + gx;
        const float cf = cz[iz] * cy[iy] * cxr[ix];
        if (REAL) {
            float *outs[3] = {o0, o1, o2};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float v = fine[fi * NC + c].x * cf * scale;
                outs[c][id] = accumulate ? outs[c][id] + v : v;
            }
        } else {
            cout[id] = cscale(fine[fi * NC], cf * scale);
        }
    }
}

// ------------------------------------------------------------------ point binning
// Sort key of every point: its tile (by the first kernel cell) times the
// tile's cell count plus the first cell's index inside the tile, so that a
// stable sort by key groups the points tile by tile and, inside a tile, the
// points that share a first kernel cell (and so touch the same cells).
// count[tile] collects the tile sizes for the chunk list.
template <int W, bool HALF>
__global__ __launch_bounds__(256) void k_bin_keys(const Geo *__restrict__ geo, long long P, int R, int C, int nfz,
                                                  int nfy, int nfx, PointFrame fr, BinParams bp,
                                                  unsigned *__restrict__ key, unsigned *__restrict__ count) {
    const long long S = (long long)R * C;
    const long long H = HALF ? (S + 1) / 2 : S;
    const long long total = P * H;
    const unsigned tvol = (unsigned)(bp.T[0] * bp.T[1] * bp.T[2]);
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * blockDim.x) {
        const long long p = gid / H;
        const long long idx = gid - p * H;
        const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
        double tx, ty, tz;
        point_coords(geo[p], fr, j, k, tx, ty, tz);
        const int ix = wrap0(first_cell(tx, nfx, fr.rx, W), nfx);
        const int iy = wrap0(first_cell(ty, nfy, fr.ry, W), nfy);
        const int iz = wrap0(first_cell(tz, nfz, fr.rz, W), nfz);
        const int bz = iz / bp.T[0], by = iy / bp.T[1], bxx = ix / bp.T[2];
        const unsigned tile = (unsigned)((bz * bp.nt[1] + by) * bp.nt[2] + bxx);
        const unsigned local =
            (unsigned)(((iz - bz * bp.T[0]) * bp.T[1] + (iy - by * bp.T[1])) * bp.T[2] + (ix - bxx * bp.T[2]));
        key[gid] = tile * tvol + local;
        atomicAdd(&count[tile], 1u);
    }
}

// ------------------------------------------------------------------ bucket sort of the bins
// The sort keys above are bucket * cells + cell, with buckets (a tile, or a
// (row, tile) pair) counted into count[].  The sorted order is by bucket, then
// by cell, then by point index.  Step 1 scatters every point into its bucket's
// range (host-scanned offsets) at a position taken from an atomic cursor, so
// the order inside a bucket is arbitrary there; step 2 gives every element of
// a bucket its rank by (cell, point index), which is unique, so the result is
// the same whatever order step 1 produced.
__global__ __launch_bounds__(256) void k_bin_scatter(const unsigned *__restrict__ key, long long total, unsigned cells,
                                                     const unsigned *__restrict__ offset, unsigned *__restrict__ cursor,
                                                     unsigned long long *__restrict__ staged) {
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * blockDim.x) {
        const unsigned k = key[gid];
        const unsigned b = k / cells;   // keys reach 2^32: plain division (once per point, set-up only)
        const unsigned cell = k - b * cells;
        const unsigned pos = offset[b] + atomicAdd(&cursor[b], 1u);
        staged[pos] = ((unsigned long long)cell << 32) | (unsigned long long)gid;
    }
}

// One workgroup per bucket (blockIdx.x + first): rank = number of the bucket's
// (cell, index) keys below one's own, counted against LDS tiles of the bucket.
// Every thread runs the same loop trip counts (they depend on n only), so the
// barriers are uniform.
constexpr int kOrderTile = 1024;
__global__ __launch_bounds__(256) void k_bin_order(const unsigned long long *__restrict__ staged,
                                                   const unsigned *__restrict__ offset,
                                                   const unsigned *__restrict__ count, unsigned first,
                                                   unsigned *__restrict__ ids) {
    __shared__ unsigned long long tile[kOrderTile];
    const unsigned b = first + blockIdx.x;
    const unsigned off = offset[b], n = count[b];
    if (n == 0) return;
    if (n == 1) {
        if (threadIdx.x == 0) ids[off] = (unsigned)staged[off];
        return;
    }
    for (unsigned i0 = 0; i0 < n; i0 += blockDim.x) {
        const unsigned i = i0 + threadIdx.x;
        const unsigned long long ki = i < n ? staged[off + i] : ~0ull;
        unsigned rank = 0;
        for (unsigned j0 = 0; j0 < n; j0 += kOrderTile) {
            __syncthreads();
            for (unsigned t = threadIdx.x; t < (unsigned)kOrderTile; t += blockDim.x)
                tile[t] = j0 + t < n ? staged[off + j0 + t] : ~0ull;
            __syncthreads();
            const unsigned m = min((unsigned)kOrderTile, n - j0);
            for (unsigned t = 0; t < m; ++t) rank += tile[t] < ki ? 1u : 0u;
        }
        if (i < n) ids[off + rank] = (unsigned)ki;
    }
}

// ------------------------------------------------------------------ tiled kernels
// A workgroup handles one chunk of the points binned to a tile (tz,ty,tx).
// The tile's halo box covers the fine cells [o, o + T + W - 1) per axis, so
// every kernel cell of its points lies inside it.
struct TileBox {
    int oz, oy, ox, HZ, HY, HX, ncell;
};

template <int W>
__device__ __forceinline__ TileBox tile_box(int tile, const BinParams &bp) {
    TileBox b;
    const int tx = tile % bp.nt[2], ty = (tile / bp.nt[2]) % bp.nt[1], tz = tile / (bp.nt[2] * bp.nt[1]);
    b.oz = tz * bp.T[0]; b.oy = ty * bp.T[1]; b.ox = tx * bp.T[2];
    b.HZ = bp.T[0] + W - 1; b.HY = bp.T[1] + W - 1; b.HX = bp.T[2] + W - 1;
    b.ncell = b.HZ * b.HY * b.HX;
    return b;
}

__device__ __forceinline__ int wrap_up(int i, int n) {
    while (i >= n) i -= n;   // a box reaches at most a few periods past its origin
    return i;
}

__device__ __forceinline__ long long box_to_fine(const TileBox &b, const BinParams &bp, int cell, int nfz, int nfy,
                                                 int nfx) {
    const int cz_ = (int)bp.fhyx.div((unsigned)cell), rem = cell - cz_ * b.HY * b.HX;
    const int cy_ = (int)bp.fhx.div((unsigned)rem), cx_ = rem - cy_ * b.HX;
    const int gz = wrap_up(b.oz + cz_, nfz), gy = wrap_up(b.oy + cy_, nfy), gx = wrap_up(b.ox + cx_, nfx);
    return ((long long)gz * nfy + gy) * nfx + gx;
}

// The three component values a half-set point spreads: v = c(q) + conj c(-q)
// of plane p (the mirror is q's negative), times d_p.  Paired planes hold
// Z = P2m + i P2m+1 in plane 2m: the even plane's value is Z(q) + conj Z(-q),
// the odd one's -i (Z(q) - conj Z(-q)); the centre point (its own mirror)
// takes Re Z or Im Z.
// a = c(q) and b = c(-q) of the point's (paired) plane, as loaded by half_point_values
__device__ __forceinline__ void half_point_combine(float2 a, float2 b, long long p, long long idx, long long S,
                                                   const Geo &g, int paired, float2 vc[3]) {
    const long long mir = S - 1 - idx;
    float2 v;
    if (!paired) v = mir != idx ? cadd(a, cconj(b)) : a;
    else if (!(p & 1)) v = mir != idx ? make_float2(a.x + b.x, a.y - b.y) : make_float2(a.x, 0.f);
    else v = mir != idx ? make_float2(a.y + b.y, b.x - a.x) : make_float2(a.y, 0.f);
    vc[0] = cscale(v, (float)g.d0);
    vc[1] = cscale(v, (float)g.d1);
    vc[2] = cscale(v, (float)g.d2);
}

__device__ __forceinline__ void half_point_values(const float2 *__restrict__ c, long long p, long long idx,
                                                  long long S, const Geo &g, int paired, float2 vc[3]) {
    const long long mir = S - 1 - idx;
    const long long pl = paired ? (p & ~1ll) : p;
    const float2 a = c[pl * S + idx];
    const float2 b = mir != idx ? c[pl * S + mir] : make_float2(0.f, 0.f);
    half_point_combine(a, b, p, idx, S, g, paired, vc);
}

// The interpolated value of a half-set point and its mirror: s = sum_i d_i c_i
// at q, conj(s) at -q; at q = 0 (its own mirror) the transform of a real
// volume is real.
__device__ __forceinline__ void store_half_point(float2 *__restrict__ c, long long p, long long idx, long long S,
                                                 const Geo &g, const float2 acc[3]) {
    const float d0 = (float)g.d0, d1 = (float)g.d1, d2 = (float)g.d2;
    float2 s;
    s.x = d0 * acc[0].x + d1 * acc[1].x + d2 * acc[2].x;
    s.y = d0 * acc[0].y + d1 * acc[1].y + d2 * acc[2].y;
    const long long mir = S - 1 - idx;
    if (mir == idx) s.y = 0.f;
    c[p * S + idx] = s;
    if (mir != idx) c[p * S + mir] = cconj(s);
}

// ---- spreading (type 1)
// The box accumulator lives in LDS as NC complex planes.  The four waves of
// the workgroup own the four (z, y) parity classes of box cells; every wave
// walks all staged points and adds the contributions to its own cells with
// plain read-modify-writes, so no cell is ever touched by two lanes or two
// waves at once and no atomics are needed.  The box is then flushed once to
// the fine grid with global atomics (neighbouring boxes overlap).
//
// The walk is bound by LDS traffic.  The points of a chunk arrive sorted by
// first kernel cell; a run of points with the same first cell covers the
// same cells, so each wave sums the run in registers and does the box
// read-modify-write once per run instead of once per point.  A point's staged
// record is laid out so that each lane fetches what it needs (header, values,
// and both parity candidates of its z and y weights) with a few aligned reads
// that do not depend on one another, and the next point's record is fetched
// while the current point is summed.
constexpr int kStage = 128;   // points staged per batch

// Staged point record, in 32-bit words: [0] box cell of the point's first
// kernel cell, [1] parity bits (lz & 1) | (ly & 1) << 1, [4, 4 + 2 NC) the NC
// complex values, then wz[W], wy[W], wx[W].  Records are 16-byte aligned.
template <int W, int NC>
struct StageLayout {
    static constexpr int VQ = 4;
    static constexpr int WZ = VQ + 2 * NC;
    static constexpr int WY = WZ + W;
    static constexpr int WX = WY + W;
    static constexpr int SW = (WX + W + 3) / 4 * 4;
    static_assert(W % 2 == 0 && WZ % 2 == 0, "weight pairs must be 8-byte aligned");
};

template <int W, int NC>
constexpr int stage_words() { return StageLayout<W, NC>::SW; }

template <int W, bool HALF>
__global__ __launch_bounds__(256) void k_spread_tiled(float2 *__restrict__ fine, const float2 *__restrict__ c,
                                                      const Geo *__restrict__ geo, long long P, int R, int C,
                                                      int nfz, int nfy, int nfx, float beta, PointFrame fr,
                                                      BinParams bp, const unsigned *__restrict__ ids,
                                                      const uint4 *__restrict__ chunks, int paired) {
    constexpr int NC = HALF ? 3 : 1;
    using SL = StageLayout<W, NC>;
    constexpr int SW = SL::SW;
    constexpr int HW = W / 2;
    constexpr int NOWN = HW * HW * W;     // cells of one point owned by one wave
    constexpr int NCH = (NOWN + 63) / 64; // lane passes over them
    extern __shared__ float lds[];
    const uint4 ch = chunks[blockIdx.x];
    const TileBox bx = tile_box<W>((int)ch.x, bp);
    float *stg = lds;                                                  // [kStage][SW], 16-byte records
    float2 *acc = reinterpret_cast<float2 *>(lds + (size_t)kStage * SW);  // [NC][ncell]
    for (int i = threadIdx.x; i < NC * bx.ncell; i += blockDim.x) acc[i] = make_float2(0.f, 0.f);

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int zpar = wave >> 1, ypar = wave & 1;
    const int HYX = bx.HY * bx.HX;
    const long long S = (long long)R * C;
    const long long H = HALF ? (S + 1) / 2 : S;

    // this lane's cells of a point: (a, b, e) = (zoff + 2 ia, yoff + 2 ib, e)
    int loff[NCH], wzo[NCH], wyo[NCH], wxo[NCH];
    bool act[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int cidx = lane + 64 * c;
        act[c] = cidx < NOWN;
        const int cc = act[c] ? cidx : 0;
        const int e = cc % W, ib = (cc / W) % HW, ia = cc / (W * HW);
        loff[c] = (2 * ia * bx.HY + 2 * ib) * bx.HX + e;
        wzo[c] = SL::WZ + 2 * ia;
        wyo[c] = SL::WY + 2 * ib;
        wxo[c] = SL::WX + e;
    }

    // The points of a chunk come sorted by first kernel cell, so runs of points
    // share one: a run is summed in registers and added to the box once.
    float2 run[NCH][NC];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int q = 0; q < NC; ++q) run[c][q] = make_float2(0.f, 0.f);
    int run_cell = -1;         // box cell of the run's first kernel cell (uniform), -1 before the first
    float2 *run_acc = acc;     // the wave's first owned cell of that run
    auto add_run = [&]() {
        if (run_cell < 0) return;
        float2 t[NCH][NC];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (act[c]) {
#pragma unroll
                for (int q = 0; q < NC; ++q) t[c][q] = run_acc[q * bx.ncell + loff[c]];
            }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (act[c]) {
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    t[c][q].x += run[c][q].x;
                    t[c][q].y += run[c][q].y;
                    run_acc[q * bx.ncell + loff[c]] = t[c][q];
                }
            }
#pragma unroll
            for (int q = 0; q < NC; ++q) run[c][q] = make_float2(0.f, 0.f);
        }
    };

    for (unsigned base = 0; base < ch.z; base += kStage) {
        const int nb = (int)min((unsigned)kStage, ch.z - base);
        __syncthreads();  // the box is zeroed / the previous batch is consumed
        if ((int)threadIdx.x < nb) {
            const long long gid = ids[ch.y + base + threadIdx.x];
            const long long p = gid / H;
            const long long idx = gid - p * H;
            const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
            const Geo g = geo[p];
            float2 vc[NC];
            if constexpr (HALF) half_point_values(c, p, idx, S, g, paired, vc);
            else vc[0] = c[p * S + idx];
            Placed<W> pt;
            place<W>(g, fr, j, k, nfz, nfy, nfx, beta, pt);
            const int lz = wrap0(pt.iz, nfz) - bx.oz, ly = wrap0(pt.iy, nfy) - bx.oy, lx = wrap0(pt.ix, nfx) - bx.ox;
            const bool inside = lz >= 0 && lz < bp.T[0] && ly >= 0 && ly < bp.T[1] && lx >= 0 && lx < bp.T[2];
            if (!inside) {
                // a point outside its tile (not expected): spread it straight to the fine grid
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    const int zz = wrapi(pt.iz + a, nfz);
#pragma unroll
                    for (int b = 0; b < W; ++b) {
                        const int yy = wrapi(pt.iy + b, nfy);
                        float2 *row = fine + ((long long)zz * nfy + yy) * (long long)nfx * NC;
                        const float wzy = pt.wz[a] * pt.wy[b];
#pragma unroll
                        for (int e = 0; e < W; ++e) {
                            float *dst = reinterpret_cast<float *>(row + (long long)wrapi(pt.ix + e, nfx) * NC);
                            const float wgt = wzy * pt.wx[e];
#pragma unroll
                            for (int q = 0; q < NC; ++q) {
                                atomicAdd(dst + 2 * q, vc[q].x * wgt);
                                atomicAdd(dst + 2 * q + 1, vc[q].y * wgt);
                            }
                        }
                    }
                }
            }
            // a point spread directly above stays in the walk as a zero record
            float *sp = stg + threadIdx.x * SW;
            int *si = reinterpret_cast<int *>(sp);
            si[0] = inside ? (lz * bx.HY + ly) * bx.HX + lx : 0;
            si[1] = inside ? (lz & 1) | ((ly & 1) << 1) : 0;
#pragma unroll
            for (int m = 0; m < W; ++m) {
                sp[SL::WZ + m] = inside ? pt.wz[m] : 0.f;
                sp[SL::WY + m] = inside ? pt.wy[m] : 0.f;
                sp[SL::WX + m] = inside ? pt.wx[m] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                sp[SL::VQ + 2 * q] = inside ? vc[q].x : 0.f;
                sp[SL::VQ + 2 * q + 1] = inside ? vc[q].y : 0.f;
            }
        }
        __syncthreads();

        // one point's record as this lane needs it
        struct Rec {
            int2 hdr;
            float2 v[NC];
            float2 wz[NCH], wy[NCH];   // parity candidates (w[2i], w[2i + 1])
            float wx[NCH];
        };
        auto fetch = [&](int kk, Rec &r) {
            const float *sp = stg + kk * SW;
            r.hdr = *reinterpret_cast<const int2 *>(sp);
#pragma unroll
            for (int q = 0; q < NC; ++q) r.v[q] = *reinterpret_cast<const float2 *>(sp + SL::VQ + 2 * q);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                r.wz[c] = *reinterpret_cast<const float2 *>(sp + wzo[c]);
                r.wy[c] = *reinterpret_cast<const float2 *>(sp + wyo[c]);
                r.wx[c] = sp[wxo[c]];
            }
        };
        auto accumulate = [&](const Rec &r) {
            const int cb = __builtin_amdgcn_readfirstlane(r.hdr.x);
            const int bits = __builtin_amdgcn_readfirstlane(r.hdr.y);
            const int zoff = (zpar ^ bits) & 1, yoff = (ypar ^ (bits >> 1)) & 1;
            if (cb != run_cell) {
                add_run();
                run_cell = cb;
                run_acc = acc + cb + (zoff ? HYX : 0) + (yoff ? bx.HX : 0);
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float w = (zoff ? r.wz[c].y : r.wz[c].x) * (yoff ? r.wy[c].y : r.wy[c].x) * r.wx[c];
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    run[c][q].x += r.v[q].x * w;
                    run[c][q].y += r.v[q].y * w;
                }
            }
        };
        // two records in flight: the next one is fetched before the current one is accumulated
        Rec ra, rb;
        fetch(0, ra);
        for (int kk = 0; kk < nb; kk += 2) {
            if (kk + 1 < nb) fetch(kk + 1, rb);
            accumulate(ra);
            if (kk + 1 < nb) {
                if (kk + 2 < nb) fetch(kk + 2, ra);
                accumulate(rb);
            }
        }
    }
    add_run();
    __syncthreads();
    for (int cell = threadIdx.x; cell < bx.ncell; cell += blockDim.x) {
        float *dst = reinterpret_cast<float *>(fine + box_to_fine(bx, bp, cell, nfz, nfy, nfx) * NC);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const float2 t = acc[q * bx.ncell + cell];
            if (t.x != 0.f) atomicAdd(dst + 2 * q, t.x);
            if (t.y != 0.f) atomicAdd(dst + 2 * q + 1, t.y);
        }
    }
}

// ---- interpolation (type 2), tiled: the tile's halo box of the fine grid is
// staged in LDS once and every point of the chunk gathers from it.
template <int W, bool HALF>
__global__ __launch_bounds__(256) void k_interp_tiled(const float2 *__restrict__ fine, float2 *__restrict__ c,
                                                      const Geo *__restrict__ geo, long long P, int R, int C,
                                                      int nfz, int nfy, int nfx, float beta, PointFrame fr,
                                                      BinParams bp, const unsigned *__restrict__ ids,
                                                      const uint4 *__restrict__ chunks) {
    constexpr int NC = HALF ? 3 : 1;
    extern __shared__ float lds[];
    const uint4 ch = chunks[blockIdx.x];
    const TileBox bx = tile_box<W>((int)ch.x, bp);
    float2 *box = reinterpret_cast<float2 *>(lds);   // [ncell][NC]
    // each thread issues the loads of kBoxBatch box values before writing them to LDS
    constexpr int kBoxBatch = 8;
    const int nbox = NC * bx.ncell;
    const int bd = blockDim.x;
    for (int i0 = threadIdx.x; i0 < nbox; i0 += kBoxBatch * bd) {
        float2 v[kBoxBatch];
#pragma unroll
        for (int u = 0; u < kBoxBatch; ++u) {
            const int i = i0 + u * bd;
            if (i < nbox) {
                const int cell = i / NC, q = i - cell * NC;
                v[u] = fine[box_to_fine(bx, bp, cell, nfz, nfy, nfx) * NC + q];
            }
        }
#pragma unroll
        for (int u = 0; u < kBoxBatch; ++u) {
            const int i = i0 + u * bd;
            if (i < nbox) box[i] = v[u];
        }
    }
    __syncthreads();
    const long long S = (long long)R * C;
    const long long H = HALF ? (S + 1) / 2 : S;
    for (unsigned i = threadIdx.x; i < ch.z; i += blockDim.x) {
        const long long gid = ids[ch.y + i];
        const long long p = gid / H;
        const long long idx = gid - p * H;
        const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
        const Geo g = geo[p];
        Placed<W> pt;
        place<W>(g, fr, j, k, nfz, nfy, nfx, beta, pt);
        const int lz = wrap0(pt.iz, nfz) - bx.oz, ly = wrap0(pt.iy, nfy) - bx.oy, lx = wrap0(pt.ix, nfx) - bx.ox;
        float2 acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = make_float2(0.f, 0.f);
        if (lz >= 0 && lz < bp.T[0] && ly >= 0 && ly < bp.T[1] && lx >= 0 && lx < bp.T[2]) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const float2 *row = box + (size_t)(((lz + a) * bx.HY + (ly + b)) * bx.HX + lx) * NC;
                    float2 racc[NC];
#pragma unroll
                    for (int q = 0; q < NC; ++q) racc[q] = make_float2(0.f, 0.f);
#pragma unroll
                    for (int e = 0; e < W; ++e) {
#pragma unroll
                        for (int q = 0; q < NC; ++q) {
                            const float2 v = row[e * NC + q];
                            racc[q].x += v.x * pt.wx[e];
                            racc[q].y += v.y * pt.wx[e];
                        }
                    }
                    const float wzy = pt.wz[a] * pt.wy[b];
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        acc[q].x += racc[q].x * wzy;
                        acc[q].y += racc[q].y * wzy;
                    }
                }
            }
        } else {
            // a point outside its tile (not expected): gather from the fine grid
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const int zz = wrapi(pt.iz + a, nfz);
#pragma unroll
                for (int b = 0; b < W; ++b) {
                    const int yy = wrapi(pt.iy + b, nfy);
                    const float2 *row = fine + ((long long)zz * nfy + yy) * (long long)nfx * NC;
                    const float wzy = pt.wz[a] * pt.wy[b];
#pragma unroll
                    for (int e = 0; e < W; ++e) {
                        const float wgt = wzy * pt.wx[e];
#pragma unroll
                        for (int q = 0; q < NC; ++q) {
                            const float2 v = row[(long long)wrapi(pt.ix + e, nfx) * NC + q];
                            acc[q].x += v.x * wgt;
                            acc[q].y += v.y * wgt;
                        }
                    }
                }
            }
        }
        if constexpr (HALF) {
            store_half_point(c, p, idx, S, g, acc);
        } else {
            c[p * S + idx] = acc[0];
        }
    }
}

// ------------------------------------------------------------------ row-factorised path
// (see RowPlan).  Row j of the half point set has the unwrapped first y cell
// iy[j] and the weights wy[j][0..W); its slab holds the (z, x) plane of
// [nfz][nfx][3] complex values.  Rows are processed in groups [j0, j0 + nrows).

__device__ __forceinline__ int wrap_any(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// Sort key of a half point for the row path: (row, 2-D tile of its (z, x)
// first kernel cell, that cell inside the tile).
template <int W>
__global__ __launch_bounds__(256) void k_bin_keys_rows(const Geo *__restrict__ geo, long long P, int R, int C,
                                                       int nfz, int nfx, PointFrame fr, RowGeom rg,
                                                       unsigned *__restrict__ key, unsigned *__restrict__ count) {
    const long long S = (long long)R * C, H = (S + 1) / 2, total = P * H;
    const unsigned tt = (unsigned)(rg.Tz * rg.Tx);
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < total;
         gid += (long long)gridDim.x * blockDim.x) {
        const long long p = gid / H;
        const long long idx = gid - p * H;
        const int j = (int)(idx / C), k = (int)(idx - (long long)j * C);
        double tx, ty, tz;
        point_coords(geo[p], fr, j, k, tx, ty, tz);
        const int ix = wrap0(first_cell(tx, nfx, fr.rx, W), nfx);
        const int iz = wrap0(first_cell(tz, nfz, fr.rz, W), nfz);
        const int bz = iz / rg.Tz, bx = ix / rg.Tx;
        const unsigned jt = (unsigned)j * rg.ntiles + (unsigned)(bz * rg.ntx + bx);
        key[gid] = jt * tt + (unsigned)((iz - bz * rg.Tz) * rg.Tx + (ix - bx * rg.Tx));
        atomicAdd(&count[jt], 1u);
    }
}

// The 2-D boxes that reach fine-grid cell (z, x): its own tile's, and for a cell
// in the first W - 1 rows / columns of its tile the boxes of the tiles above
// (S), to the left (E) and above-left (C), which extend W - 1 cells past their
// tile.  A row's slab value at (z, x) is read (interpolation) or written
// (spreading) only through a box of a tile that holds points of that row.
struct CellTiles {
    unsigned t0, tS, tE, tC;
    int lz, lx;
    bool south, east;
};

template <int W>
__device__ __forceinline__ CellTiles cell_tiles(int z, int x, const RowGeom &rg) {
    CellTiles c;
    const int tz = z / rg.Tz, tx = x / rg.Tx;
    c.lz = z - tz * rg.Tz;
    c.lx = x - tx * rg.Tx;
    const int tzu = tz > 0 ? tz - 1 : rg.ntz - 1, txl = tx > 0 ? tx - 1 : rg.ntx - 1;
    c.t0 = (unsigned)(tz * rg.ntx + tx);
    c.tS = (unsigned)(tzu * rg.ntx + tx);
    c.tE = (unsigned)(tz * rg.ntx + txl);
    c.tC = (unsigned)(tzu * rg.ntx + txl);
    c.south = c.lz < W - 1;
    c.east = c.lx < W - 1;
    return c;
}

// Which of those boxes hold points in row j, from the plan's bit table ([tile][word],
// bit j % 32 of word j / 32).  The rows are walked in one direction (dir), so the
// four words of the next 32-row block are loaded while the current block is used.
struct NzRow {
    int blk = -1;
    unsigned b0 = 0, bS = 0, bE = 0, bC = 0;      // current block
    unsigned q0 = 0, qS = 0, qE = 0, qC = 0;      // next block (in the walk's direction)
    bool n0 = false, nS = false, nE = false, nC = false;
    __device__ __forceinline__ void load(const unsigned *__restrict__ nz, int nzw, const CellTiles &c, int b,
                                         unsigned &w0, unsigned &wS, unsigned &wE, unsigned &wC) {
        b = b < 0 ? 0 : (b >= nzw ? nzw - 1 : b);
        w0 = nz[(long long)c.t0 * nzw + b];
        wS = c.south ? nz[(long long)c.tS * nzw + b] : 0u;
        wE = c.east ? nz[(long long)c.tE * nzw + b] : 0u;
        wC = c.south && c.east ? nz[(long long)c.tC * nzw + b] : 0u;
    }
    __device__ __forceinline__ void at(const unsigned *__restrict__ nz, int nzw, const CellTiles &c, int j,
                                       int dir) {
        const int b = j >> 5;
        if (b != blk) {
            if (blk < 0) load(nz, nzw, c, b, q0, qS, qE, qC);
            b0 = q0; bS = qS; bE = qE; bC = qC;
            blk = b;
            load(nz, nzw, c, b + (dir > 0 ? 1 : -1), q0, qS, qE, qC);
        }
        const unsigned sh = (unsigned)(j & 31);
        n0 = (b0 >> sh) & 1u;
        nS = (bS >> sh) & 1u;
        nE = (bE >> sh) & 1u;
        nC = (bC >> sh) & 1u;
    }
    __device__ __forceinline__ bool any() const { return n0 || nS || nE || nC; }
};

// type 2, step 1: G_r(z, x) = sum_b wy[j][b] F(z, iy[j] + b, x), j = j0 + r.  One thread per
// (z, x, component) walks the group's rows in the order of non-decreasing iy with
// a window of W fine-grid values in registers, so each value is loaded once.
template <int W>
__global__ __launch_bounds__(256) void k_rows_contract(const float2 *__restrict__ fine, float2 *__restrict__ slab,
                                                       const float *__restrict__ wy, const int *__restrict__ iy,
                                                       int nfz, int nfy, int nfx, int j0, int nrows, int dir) {
    const long long nxc = (long long)nfx * 3, total = (long long)nfz * nxc;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(id / nxc);
        const long long xc = id - (long long)z * nxc;
        const float2 *col = fine + (long long)z * nfy * nxc + xc;
        int ybase = iy[j0 + (dir > 0 ? 0 : nrows - 1)];
        float2 win[W];
#pragma unroll
        for (int b = 0; b < W; ++b) win[b] = col[(long long)wrap_any(ybase + b, nfy) * nxc];
        for (int t = 0; t < nrows; ++t) {
            const int r = dir > 0 ? t : nrows - 1 - t;
            const int j = j0 + r;
            const int iyr = iy[j];
            while (ybase < iyr) {
#pragma unroll
                for (int b = 0; b < W - 1; ++b) win[b] = win[b + 1];
                win[W - 1] = col[(long long)wrap_any(ybase + W, nfy) * nxc];
                ++ybase;
            }
            const float *w = wy + (long long)j * W;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int b = 0; b < W; ++b) {
                acc.x += w[b] * win[b].x;
                acc.y += w[b] * win[b].y;
            }
            slab[((long long)r * nfz + z) * nxc + xc] = acc;
        }
    }
}

// type 1, step 2: F(z, iy[j] + b, x) += wy[j][b] H_r(z, x) over the group's rows.  H_r(z, x)
// is the interior value its tile's workgroup stored plus the halo strips of the
// tiles above, left and above-left whose boxes reach (z, x); tiles without points
// stored nothing and count as zero.  One thread per (z, x, component) owns that
// fine-grid column (no atomics), with a window of W accumulators in registers.
// The thread writes each plane of its y window [iy_first, iy_last + W) once;
// accumulate = 0 stores instead of adding (one row group whose window does not
// wrap: the planes need no zeroing and the flushes wait on no load).
template <int W>
__global__ __launch_bounds__(256) void k_rows_expand(float2 *__restrict__ fine, const float2 *__restrict__ slab,
                                                     const float2 *__restrict__ halo, const unsigned *__restrict__ nz,
                                                     int nzw, const float *__restrict__ wy, const int *__restrict__ iy,
                                                     int nfz, int nfy, int nfx, RowGeom rg, int j0, int nrows,
                                                     int dir, int accumulate) {
    const long long nxc = (long long)nfx * 3, total = (long long)nfz * nxc;
    for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(id / nxc);
        const long long xc = id - (long long)z * nxc;
        const int x = (int)(xc / 3), m = (int)(xc - 3ll * x);
        const CellTiles ct = cell_tiles<W>(z, x, rg);
        const int lz = ct.lz, lx = ct.lx;
        const int hS = lz * rg.Tx + lx, hE = (W - 1) * rg.Tx + lz * (W - 1) + lx,
                  hC = (W - 1) * rg.Tx + rg.Tz * (W - 1) + lz * (W - 1) + lx;
        NzRow nr;
        float2 *col = fine + (long long)z * nfy * nxc + xc;
        int ybase = iy[j0 + (dir > 0 ? 0 : nrows - 1)];
        float2 acc[W];
#pragma unroll
        for (int b = 0; b < W; ++b) acc[b] = make_float2(0.f, 0.f);
        // A row's slab and halo loads are issued before the window planes it closes are
        // written, so the two overlap (measured against loading after the flush: 1.74 ->
        // 1.45 ms per launch at (359, 71, 359) / 615; loading two or four rows at once
        // gained nothing more).  Only the tiles whose boxes hold points of the row stored
        // values there; the others' cells are not read (a quarter of the (row, tile)
        // pairs of a polar grid are empty), and a value not loaded adds as an exact zero.
        for (int t = 0; t < nrows; ++t) {
            const int r = dir > 0 ? t : nrows - 1 - t;
            const int j = j0 + r;
            nr.at(nz, nzw, ct, j, dir);
            const float2 *hr = halo + (long long)r * rg.ntiles * rg.NH * 3;
            float2 v0 = make_float2(0.f, 0.f), vS = v0, vE = v0, vC = v0;
            if (nr.n0) v0 = slab[((long long)r * nfz + z) * nxc + xc];
            if (nr.nS) vS = hr[((long long)ct.tS * rg.NH + hS) * 3 + m];
            if (nr.nE) vE = hr[((long long)ct.tE * rg.NH + hE) * 3 + m];
            if (nr.nC) vC = hr[((long long)ct.tC * rg.NH + hC) * 3 + m];
            const int iyr = iy[j];
            while (ybase < iyr) {
                float2 *dst = col + (long long)wrap_any(ybase, nfy) * nxc;
                *dst = accumulate ? cadd(*dst, acc[0]) : acc[0];
#pragma unroll
                for (int b = 0; b < W - 1; ++b) acc[b] = acc[b + 1];
                acc[W - 1] = make_float2(0.f, 0.f);
                ++ybase;
            }
            if (!nr.any()) continue;
            const float2 h = cadd(cadd(cadd(v0, vS), vE), vC);
            const float *w = wy + (long long)j * W;
#pragma unroll
            for (int b = 0; b < W; ++b) {
                acc[b].x += w[b] * h.x;
                acc[b].y += w[b] * h.y;
            }
        }
#pragma unroll
        for (int b = 0; b < W; ++b) {
            float2 *dst = col + (long long)wrap_any(ybase + b, nfy) * nxc;
            *dst = accumulate ? cadd(*dst, acc[b]) : acc[b];
        }
    }
}

// Windowed form of the contraction, for groups whose rows span few y planes.
// A workgroup owns blockDim.x fine-grid columns; the group's y window of each
// column (span planes from lo, the first processed row's first cell) is loaded
// into LDS at once as [span][blockDim.x], each thread touching only its own
// column (so there are no barriers), and the row walk then waits on LDS instead
// of on one dependent global load per y step (0.164 -> 0.130 ms per launch at
// the benchmark shape).  The same window for the expansion measured slower
// (0.30 -> 0.40 ms: its 64 KiB of LDS per workgroup leaves 8 waves per CU
// against 20), and so did issuing the expansion's next-row loads a row ahead
// (0.30 -> 0.32 ms).
template <int W>
__global__ __launch_bounds__(256) void k_rows_contract_win(const float2 *__restrict__ fine, float2 *__restrict__ slab,
                                                           const float *__restrict__ wy, const int *__restrict__ iy,
                                                           int nfz, int nfy, int nfx, int j0, int nrows, int dir,
                                                           int lo, int span) {
    extern __shared__ float2 wl[];
    const int bd = blockDim.x, tid = threadIdx.x;
    const long long nxc = (long long)nfx * 3, total = (long long)nfz * nxc;
    const long long id = (long long)blockIdx.x * bd + tid;
    if (id >= total) return;
    const int z = (int)(id / nxc);
    const long long xc = id - (long long)z * nxc;
    const float2 *col = fine + (long long)z * nfy * nxc + xc;
    float2 *wc = wl + tid;
#pragma unroll 16
    for (int s = 0; s < span; ++s) wc[s * bd] = col[(long long)wrap_any(lo + s, nfy) * nxc];
    float2 win[W];
#pragma unroll
    for (int b = 0; b < W; ++b) win[b] = wc[b * bd];
    int yb = 0;
    for (int t = 0; t < nrows; ++t) {
        const int r = dir > 0 ? t : nrows - 1 - t;
        const int j = j0 + r;
        const int o = iy[j] - lo;
        while (yb < o) {   // yb + W <= o - 1 + W <= span - 1
#pragma unroll
            for (int b = 0; b < W - 1; ++b) win[b] = win[b + 1];
            win[W - 1] = wc[(yb + W) * bd];
            ++yb;
        }
        const float *wj = wy + (long long)j * W;
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int b = 0; b < W; ++b) {
            acc.x += wj[b] * win[b].x;
            acc.y += wj[b] * win[b].y;
        }
        slab[((long long)r * nfz + z) * nxc + xc] = acc;
    }
}

// Box of a 2-D tile: cells [o, o + HB) per axis.
struct Box2 {
    int r, tile, oz, ox, ncell;
};

__device__ __forceinline__ Box2 box2(unsigned jt, int j0, const RowGeom &rg) {
    Box2 b;
    const int j = (int)(jt / rg.ntiles);
    b.tile = (int)(jt - (unsigned)j * rg.ntiles);
    const int tz = b.tile / rg.ntx;
    b.r = j - j0;
    b.oz = tz * rg.Tz;
    b.ox = (b.tile - tz * rg.ntx) * rg.Tx;
    b.ncell = rg.HBz * rg.HBx;
    return b;
}

// Points per (row, tile) average ~140 at the benchmark shape: 128 threads per
// workgroup (measured: 0.47 ms per launch at 256 threads, 0.43 at 128).  The
// occupancy hint has the compiler fit the kernel in fewer registers (W = 6: 63
// VGPRs instead of 112, 8 waves per SIMD, which the LDS allows): 0.343 -> 0.290
// ms.  Width 4 takes a hint of 5 and widths from 12 none, where 6 would spill.
constexpr int kRowsInterpThreads = 128;
constexpr int rows_interp_waves(int W) { return W == 4 ? 5 : W >= 12 ? 1 : 6; }

// type 2, step 2: the chunk's 2-D box of its row's slab is staged in LDS
// ([ncell][3]); each point gathers W^2 x 3 values and writes sum_i d_i c_i and
// its mirror.
template <int W>
__global__ __launch_bounds__(kRowsInterpThreads, rows_interp_waves(W)) void k_rows_interp(const float2 *__restrict__ slab, float2 *__restrict__ c,
                                                     const Geo *__restrict__ geo, long long P, int R, int C,
                                                     int nfz, int nfx, float beta, PointFrame fr, RowGeom rg,
                                                     int j0, FastDiv fH, FastDiv fC,
                                                     const unsigned *__restrict__ ids,
                                                     const uint4 *__restrict__ chunks) {
    extern __shared__ float lds[];
    const uint4 ch = chunks[blockIdx.x];
    const Box2 bx = box2(ch.x, j0, rg);
    const float2 *rs = slab + (long long)bx.r * nfz * nfx * 3;
    float2 *box = reinterpret_cast<float2 *>(lds);
    // each thread issues the loads of kBoxBatch box values before writing them to
    // LDS (one at a time, the staging waited on each load in turn: 0.43 -> 0.35 ms
    // per launch at the benchmark shape)
    // Box element i = (cz * HBx + cx) * 3 + q.  A box reaches at most W - 2 cells past the
    // end of an axis (tile edges divide the axis and are at least W), so one conditional
    // subtraction wraps it; offsets inside one row's slab fit 32 bits.
    constexpr int kBoxBatch = 8;
    const int nbox = 3 * bx.ncell, rw = 3 * rg.HBx;
    for (int i0 = threadIdx.x; i0 < nbox; i0 += kBoxBatch * kRowsInterpThreads) {
        float2 v[kBoxBatch];
#pragma unroll
        for (int u = 0; u < kBoxBatch; ++u) {
            const int i = i0 + u * kRowsInterpThreads;
            if (i < nbox) {
                const int cz = (int)rg.fbx3.div((unsigned)i), e = i - cz * rw;
                const int cxx = (int)((unsigned)e / 3u), q = e - 3 * cxx;
                int gz = bx.oz + cz, gx = bx.ox + cxx;
                gz -= gz >= nfz ? nfz : 0;
                gx -= gx >= nfx ? nfx : 0;
                v[u] = rs[(gz * nfx + gx) * 3 + q];
            }
        }
#pragma unroll
        for (int u = 0; u < kBoxBatch; ++u) {
            const int i = i0 + u * kRowsInterpThreads;
            if (i < nbox) box[i] = v[u];
        }
    }
    __syncthreads();
    const long long S = (long long)R * C;
    for (unsigned i = threadIdx.x; i < ch.z; i += blockDim.x) {
        const unsigned gid = ids[ch.y + i];
        const unsigned pu = fH.div(gid);
        const long long p = pu, idx = gid - pu * fH.d;
        const unsigned ju = fC.div((unsigned)idx);
        const int j = (int)ju, k = (int)((unsigned)idx - ju * fC.d);
        const Geo g = geo[p];
        double tx, ty, tz;
        point_coords(g, fr, j, k, tx, ty, tz);
        const int ix = first_cell(tx, nfx, fr.rx, W), iz = first_cell(tz, nfz, fr.rz, W);
        float wz[W], wx[W];
        es_weights<W>(tz, iz, beta, wz);
        es_weights<W>(tx, ix, beta, wx);
        const int lz = wrap0(iz, nfz) - bx.oz, lx = wrap0(ix, nfx) - bx.ox;
        float2 acc[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] = make_float2(0.f, 0.f);
        if (lz >= 0 && lz < rg.Tz && lx >= 0 && lx < rg.Tx) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const float2 *row = box + (size_t)((lz + a) * rg.HBx + lx) * 3;
                float2 racc[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) racc[q] = make_float2(0.f, 0.f);
#pragma unroll
                for (int e = 0; e < W; ++e) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const float2 v = row[e * 3 + q];
                        racc[q].x += v.x * wx[e];
                        racc[q].y += v.y * wx[e];
                    }
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    acc[q].x += racc[q].x * wz[a];
                    acc[q].y += racc[q].y * wz[a];
                }
            }
        } else {
            // a point outside its tile (not expected): gather from the slab
            for (int a = 0; a < W; ++a) {
                const int zz = wrapi(iz + a, nfz);
                for (int e = 0; e < W; ++e) {
                    const float2 *v = rs + ((long long)zz * nfx + wrapi(ix + e, nfx)) * 3;
                    const float wgt = wz[a] * wx[e];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        acc[q].x += v[q].x * wgt;
                        acc[q].y += v[q].y * wgt;
                    }
                }
            }
        }
        store_half_point(c, p, idx, S, g, acc);
    }
}

// type 1, step 1: one wave per (row, tile) accumulates its 2-D box ([3][ncell]
// in LDS).  Lane l < W^2 owns cell (l / W, l % W) of a point's stamp; points
// come sorted by first kernel cell, so a run of points sharing one is summed
// in registers and added to the box once.  The box's interior goes to the
// slab and the rest to the tile's halo record, both with plain stores.
constexpr int kStageRows = 64;

// x blocks: points whose first cells share z and lie in one block of B x cells
// are summed into the same registers; the wave's lanes then cover W rows of
// W + B - 1 cells.  B = 64 / W - W + 1 for one lane pass (5 at W = 6), 1 when
// W^2 needs more than one.
constexpr int row_xblock_max(int W) { return W * W <= 64 ? 64 / W - W + 1 : 1; }

// The walk.  Staging lane k keeps its record's first cell in a register and
// compares it with lane k - 1's: a ballot gives the batch's run starts as one
// scalar mask, so the walk reads no header per record and fetches the run's
// cell (readlane) only where a run opens.  A record holds v (3 complex),
// wz[W] and wx[WB] at a stride of 2 mod 4 floats (even, so v stays 8-byte
// aligned), which spreads the staging writes of consecutive lanes over the
// LDS banks.  Batches are padded to a multiple of 4 zero records so that the
// walk is unrolled by 4 without a remainder, and records are fetched two
// ahead of the one being added.  (Measured against a walk that read a header
// per record, waited on each record's reads and zero-filled its run
// registers: 6.24 -> 5.21 ms per launch at (359, 71, 359) / 615.)
template <int W>
struct RowStage {
    static constexpr int WB = W + row_xblock_max(W) - 1;
    static constexpr int WZ = 6;
    static constexpr int WX = WZ + W;
    static constexpr int N = WX + WB;
    static constexpr int SW = N + (6 - N % 4) % 4;
};

// Chunks per spreading workgroup.  The batches of its chunks form one software
// pipeline: before the walk of batch k the wave issues the geometry and value
// loads of batch k + 1 (whose point ids it loaded one batch earlier) and the id
// load of batch k + 2, and it stages batch k + 1 right after the walk of batch k,
// before the output of a finished chunk, so that no store sits between those
// loads and their use.  Only the first batch of a workgroup waits on the whole
// chain (descriptor, ids, geometry and values).  Measured at (359, 71, 359) / 615
// with the 12 x 15 tiling: 13.36 -> 12.97 ms per projector step, results
// bit-identical; 4 chunks per workgroup (descriptors re-read by scalar loads)
// 13.10, and 1.24 -> 1.33 ms at (179, 43, 179) / 305.
constexpr unsigned kSpreadChunks = 2;   // the kernel holds their descriptors in two registers

template <int W>
__global__ __launch_bounds__(64) void k_rows_spread(float2 *__restrict__ slab, float2 *__restrict__ halo,
                                                    const float2 *__restrict__ c, const Geo *__restrict__ geo,
                                                    int R, int C, int nfz, int nfx, float beta, PointFrame fr,
                                                    RowGeom rg, int j0, FastDiv fH, FastDiv fC,
                                                    const unsigned *__restrict__ ids,
                                                    const uint4 *__restrict__ chunks, unsigned nchunks, int paired) {
    using SL = RowStage<W>;
    constexpr int SW = SL::SW;
    constexpr int NCH = (W * W + 63) / 64;
    const int XB = rg.XB;
    const int WBr = W + XB - 1;
    const int NST = W * WBr;
    extern __shared__ float lds[];
    const unsigned cfirst = blockIdx.x * kSpreadChunks;
    const unsigned nown = min(kSpreadChunks, nchunks - cfirst);   // this workgroup's chunks
    const uint4 ch0 = chunks[cfirst];
    const uint4 ch1 = nown > 1 ? chunks[cfirst + 1] : make_uint4(0u, 0u, 0u, 0u);
    auto chunk = [&](unsigned t) { return t == 0 ? ch0 : ch1; };   // no indexed array: no scratch
    const int ncell = rg.HBz * rg.HBx;
    float *stg = lds;                                                        // [kStageRows + 2][SW]
    float2 *acc = reinterpret_cast<float2 *>(lds + (size_t)(kStageRows + 2) * SW);  // [3][ncell]
    for (int i = threadIdx.x; i < 3 * ncell; i += blockDim.x) acc[i] = make_float2(0.f, 0.f);
    const int lane = threadIdx.x;
    const long long S = (long long)R * C;
    const int HB = rg.HBx;

    int aoff[NCH][3], wzo[NCH], wxo[NCH];
    bool act[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) {
        const int l = lane + 64 * q;
        act[q] = l < NST;
        const int cc = act[q] ? l : 0;
        const int a = cc / WBr, e = cc - a * WBr;
#pragma unroll
        for (int m = 0; m < 3; ++m) aoff[q][m] = m * ncell + a * HB + e;
        wzo[q] = SL::WZ + a;
        wxo[q] = SL::WX + e;
    }
    float2 run[NCH][3];
#pragma unroll
    for (int q = 0; q < NCH; ++q)
#pragma unroll
        for (int m = 0; m < 3; ++m) run[q][m] = make_float2(0.f, 0.f);
    // wave-uniform; the first run start of a chunk flushes the zero run registers into
    // cell 0, which adds exact zeros, so no test of a "no run yet" state is needed
    int run_cell = 0;
    auto flush = [&]() {
        float2 t[NCH][3];
#pragma unroll
        for (int q = 0; q < NCH; ++q)
            if (act[q]) {
#pragma unroll
                for (int m = 0; m < 3; ++m) t[q][m] = acc[aoff[q][m] + run_cell];
            }
#pragma unroll
        for (int q = 0; q < NCH; ++q)
            if (act[q]) {
#pragma unroll
                for (int m = 0; m < 3; ++m) acc[aoff[q][m] + run_cell] = cadd(t[q][m], run[q][m]);
            }
    };

    // The batch sequence: (chunk t, first point base) in chunk order.
    struct Bat {
        unsigned t, base;
    };
    auto next = [&](Bat b) {
        if (b.t >= nown) return b;   // past the last batch: stays there
        b.base += kStageRows;
        if (b.base >= chunk(b.t).z) {
            ++b.t;
            b.base = 0;
        }
        return b;
    };
    auto valid = [&](const Bat &b) { return b.t < nown; };
    // Per lane: the point id of its record of a batch, then that point's geometry and
    // plane values.  Lanes past the batch load nothing.
    struct Pts {
        unsigned gid;
        Geo g;
        float2 a, b;
    };
    auto load_ids = [&](const Bat &b, unsigned &gid) {
        gid = 0;
        if (valid(b) && b.base + lane < chunk(b.t).z) gid = ids[chunk(b.t).y + b.base + lane];
    };
    auto load_pts = [&](const Bat &b, unsigned gid, Pts &pt) {
        pt.gid = gid;
        if (valid(b) && b.base + lane < chunk(b.t).z) {
            const unsigned p = fH.div(gid);
            const long long idx = gid - p * fH.d;
            pt.g = geo[p];
            const long long mir = S - 1 - idx;
            const long long pl = paired ? (long long)(p & ~1u) : (long long)p;
            pt.a = c[pl * S + idx];
            pt.b = mir != idx ? c[pl * S + mir] : make_float2(0.f, 0.f);
        }
    };

    // Staging: lane k writes its record (values, z weights, shifted x weights) and keeps
    // its first cell; a ballot gives the batch's run starts.
    int cell = -1, nb4 = 0;
    unsigned long long starts = 0;
    auto stage = [&](const Bat &b, const Pts &pt) {
        const uint4 ch = chunk(b.t);
        const Box2 bx = box2(ch.x, j0, rg);
        const int nb = (int)min((unsigned)kStageRows, ch.z - b.base);
        nb4 = (nb + 3) & ~3;
        cell = -1;
        float *sp = stg + lane * SW;
        if (lane < nb) {
            const unsigned p = fH.div(pt.gid);
            const unsigned idx = pt.gid - p * fH.d;
            const unsigned j = fC.div(idx);
            const unsigned k = idx - j * fC.d;
            float2 vc[3];
            half_point_combine(pt.a, pt.b, (long long)p, (long long)idx, S, pt.g, paired, vc);
            double tx, ty, tz;
            point_coords(pt.g, fr, (int)j, (int)k, tx, ty, tz);
            const int ix = first_cell(tx, nfx, fr.rx, W), iz = first_cell(tz, nfz, fr.rz, W);
            float wz[W], wx[W];
            es_weights<W>(tz, iz, beta, wz);
            es_weights<W>(tx, ix, beta, wx);
            int lz = wrap0(iz, nfz) - bx.oz, lx = wrap0(ix, nfx) - bx.ox;
            // the binning placed the point in this tile with the same (uncontracted FP64)
            // arithmetic, so it lies inside; the clamp only guarantees that no LDS access
            // can leave the box
            lz = min(max(lz, 0), rg.Tz - 1);
            lx = min(max(lx, 0), rg.Tx - 1);
            const int x0 = lx - lx % XB, dx = lx - x0;   // block start; x0 + WBr <= HB as XB divides Tx
            cell = lz * HB + x0;
#pragma unroll
            for (int m = 0; m < 3; ++m) *reinterpret_cast<float2 *>(sp + 2 * m) = vc[m];
#pragma unroll
            for (int a = 0; a < W; ++a) sp[SL::WZ + a] = wz[a];
            // the x weights shifted by dx into the block's WB columns, zero elsewhere (the
            // lane's own writes land in order)
#pragma unroll
            for (int e = 0; e < SL::WB; ++e) sp[SL::WX + e] = 0.f;
#pragma unroll
            for (int m = 0; m < W; ++m) sp[SL::WX + dx + m] = wx[m];
        } else if (lane < nb4) {
            // padding records: zero values and weights (LDS may hold any bit pattern)
#pragma unroll
            for (int i = 0; i < SL::N; ++i) sp[i] = 0.f;
        }
        const int up = __shfl_up(cell, 1);
        const int prev = lane == 0 ? (b.base == 0 ? -1 : run_cell) : up;
        starts = __ballot(lane < nb && cell != prev);
    };

    // Box output of a finished chunk: the interior to the slab, the rest to the tile's
    // halo record, both with plain stores.
    auto output = [&](const uint4 &ch) {
        const Box2 bx = box2(ch.x, j0, rg);
        float2 *rs = slab + (long long)bx.r * nfz * nfx * 3;
        float2 *hrec = halo + ((long long)bx.r * rg.ntiles + bx.tile) * rg.NH * 3;
        for (int cl = threadIdx.x; cl < ncell; cl += blockDim.x) {
            const int cz = (int)rg.fbx.div((unsigned)cl), cxx = cl - cz * HB;
            float2 *dst;
            if (cz < rg.Tz && cxx < rg.Tx) {
                dst = rs + ((long long)(bx.oz + cz) * nfx + (bx.ox + cxx)) * 3;
            } else {
                int h;
                if (cz >= rg.Tz && cxx < rg.Tx) h = (cz - rg.Tz) * rg.Tx + cxx;
                else if (cz < rg.Tz) h = (W - 1) * rg.Tx + cz * (W - 1) + (cxx - rg.Tx);
                else h = (W - 1) * rg.Tx + rg.Tz * (W - 1) + (cz - rg.Tz) * (W - 1) + (cxx - rg.Tx);
                dst = hrec + (long long)h * 3;
            }
#pragma unroll
            for (int m = 0; m < 3; ++m) dst[m] = acc[m * ncell + cl];
        }
    };

    // records are fetched two ahead of the one being added (the two rows past the
    // staging area's last are fetched, never written and never used, so that the
    // fetch addresses are immediate offsets from one base per group of four)
    struct Rec {
        float2 v[3];
        float wz[NCH], wx[NCH];
    };
    auto fetch = [&](int kk, Rec &r) {
        const float *rp = stg + kk * SW;
#pragma unroll
        for (int m = 0; m < 3; ++m) r.v[m] = *reinterpret_cast<const float2 *>(rp + 2 * m);
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            r.wz[q] = rp[wzo[q]];
            r.wx[q] = rp[wxo[q]];
        }
    };
    auto add = [&](bool start, int kk, const Rec &r) {
        if (start) {
            flush();
            run_cell = __builtin_amdgcn_readlane(cell, kk);
#pragma unroll
            for (int q = 0; q < NCH; ++q)
#pragma unroll
                for (int m = 0; m < 3; ++m) run[q][m] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const float w = r.wz[q] * r.wx[q];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                run[q][m].x += r.v[m].x * w;
                run[q][m].y += r.v[m].y * w;
            }
        }
    };

    // prologue: batch 0 loaded and staged, the ids of batch 1 in flight
    Bat cur = {0u, 0u};
    Pts pt;
    unsigned gid1, gid0;
    load_ids(cur, gid0);
    load_pts(cur, gid0, pt);
    Bat b1 = next(cur);
    load_ids(b1, gid1);
    __syncthreads();   // the box is zeroed
    stage(cur, pt);
    while (valid(cur)) {
        const Bat b2 = next(b1);
        // the loads of the two batches ahead, issued before this walk
        load_pts(b1, gid1, pt);
        unsigned gid2;
        load_ids(b2, gid2);
        __syncthreads();   // the staging writes are visible
        {
            Rec r0, r1, r2, r3;
            fetch(0, r0);
            fetch(1, r1);
            for (int kk = 0; kk < nb4; kk += 4) {
                // the group's four run-start bits, tested against constants
                const unsigned st = (unsigned)(starts >> kk);
                fetch(kk + 2, r2);
                add(st & 1u, kk, r0);
                fetch(kk + 3, r3);
                add(st & 2u, kk + 1, r1);
                fetch(kk + 4, r0);
                add(st & 4u, kk + 2, r2);
                fetch(kk + 5, r1);
                add(st & 8u, kk + 3, r3);
            }
        }
        const bool chunk_done = !valid(b1) || b1.t != cur.t;
        if (chunk_done) flush();
        __syncthreads();   // the walk has read the staging rows
        if (valid(b1)) stage(b1, pt);
        if (chunk_done) {
            __syncthreads();
            output(chunk(cur.t));
            if (valid(b1)) {
                __syncthreads();
                for (int i = threadIdx.x; i < 3 * ncell; i += blockDim.x) acc[i] = make_float2(0.f, 0.f);
                run_cell = 0;
#pragma unroll
                for (int q = 0; q < NCH; ++q)
#pragma unroll
                    for (int m = 0; m < 3; ++m) run[q][m] = make_float2(0.f, 0.f);
            }
        }
        cur = b1;
        b1 = b2;
        gid1 = gid2;
    }
}

// ------------------------------------------------------------------ host helpers
size_t next_smooth_even(size_t n) {
    for (;; ++n) {
        if (n % 2) continue;
        size_t r = n;
        for (size_t p : {2, 3, 5}) while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

void gauss_legendre(int n, double *x, double *wt) {
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
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

PointFrame frame(const Grid &g, const NufftPlan &pl) {
    PointFrame f;
    f.dX = kTwoPi / (double)g.C;
    f.dY = kTwoPi / (double)g.R;
    f.k0 = 0.5 * (double)(g.C - 1);
    f.j0 = 0.5 * (double)(g.R - 1);
    f.sx = (double)pl.nf[2] / kTwoPi;
    f.sy = (double)pl.nf[1] / kTwoPi;
    f.sz = (double)pl.nf[0] / kTwoPi;
    f.rx = 1.0 / (double)pl.nf[2];
    f.ry = 1.0 / (double)pl.nf[1];
    f.rz = 1.0 / (double)pl.nf[0];
    return f;
}

#define LAM_WIDTH_SWITCH(w, CALL)                                          \
    switch (w) {                                                           \
        case 4: { constexpr int W = 4; CALL; } break;                      \
        case 6: { constexpr int W = 6; CALL; } break;                      \
        case 8: { constexpr int W = 8; CALL; } break;                      \
        case 10: { constexpr int W = 10; CALL; } break;                    \
        case 12: { constexpr int W = 12; CALL; } break;                    \
        default: throw Error(LAM_ERR_NUFFT, "Error in finufft_makeplan: unsupported kernel width"); \
    }

constexpr unsigned kChunk = 2048;  // points per spreading workgroup

constexpr size_t kLdsBudget = 80 * 1024;     // largest box any tile shape may use (2 workgroups per CU)
constexpr size_t kLdsDefault = 53 * 1024;    // default shape: 3 spreading workgroups per CU (160 KiB LDS)

size_t box_cells(const int T[3], int w) {
    return (size_t)(T[0] + w - 1) * (size_t)(T[1] + w - 1) * (size_t)(T[2] + w - 1);
}

size_t spread_lds_bytes(const int T[3], int w, int nc) {
    const size_t sw = (size_t)(4 + 2 * nc + 3 * w + 3) / 4 * 4;   // StageLayout<w, nc>::SW
    return box_cells(T, w) * nc * 8 + (size_t)kStage * sw * 4;
}

size_t interp_lds_bytes(const int T[3], int w, int nc) { return box_cells(T, w) * nc * 8; }

BinParams make_bins(const int T[3], const NufftPlan &pl) {
    BinParams bp;
    for (int a = 0; a < 3; ++a) {
        bp.T[a] = T[a];
        bp.nt[a] = (pl.nf[a] + T[a] - 1) / T[a];
    }
    const unsigned hx = (unsigned)(T[2] + pl.w - 1), hy = (unsigned)(T[1] + pl.w - 1);
    bp.fhx = FastDiv(hx);
    bp.fhyx = FastDiv(hx * hy);
    return bp;
}

std::vector<long long> tune_key(const Grid &g, const NufftPlan &pl, int nc) {
    return {pl.nf[0], pl.nf[1], pl.nf[2], pl.w, nc, g.R, g.C};
}

bool bins_fit(const int T[3], const NufftPlan &pl, int nc) {
    for (int a = 0; a < 3; ++a)
        if (T[a] < 2 || T[a] > pl.nf[a]) return false;
    return spread_lds_bytes(T, pl.w, nc) <= kLdsBudget && interp_lds_bytes(T, pl.w, nc) <= kLdsBudget;
}

// Tile shapes chosen by the autotuner, shared by every context of the process
// (a shape tuned by one server worker serves the others).
std::mutex g_tune_mu;
std::map<std::vector<long long>, std::array<int, 3>> g_tune;

bool lookup_tuned(Ctx &cx, const std::vector<long long> &key, std::array<int, 3> &T) {
    auto it = cx.bin_tune.find(key);
    if (it != cx.bin_tune.end()) {
        T = it->second;
        return true;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto jt = g_tune.find(key);
    if (jt == g_tune.end()) return false;
    T = jt->second;
    cx.bin_tune[key] = T;
    return true;
}

// The tuned tile shape of this grid and fine grid when there is one, else the
// largest cubic tile whose spreading box fits the LDS budget.
BinParams bin_params(Ctx &cx, const Grid &g, const NufftPlan &pl, int nc) {
    std::array<int, 3> tuned;
    if (lookup_tuned(cx, tune_key(g, pl, nc), tuned)) return make_bins(tuned.data(), pl);
    // Untuned: the largest cubic tile whose spreading box leaves room for three
    // workgroups per CU (measured faster than two workgroups on bigger boxes:
    // the spreader's LDS walk needs the extra waves to hide its latency).
    int T[3] = {2, 2, 2};
    for (int cand : {16, 12, 10, 8, 7, 6, 5, 4, 2}) {
        const int c3[3] = {cand, cand, cand};
        if (spread_lds_bytes(c3, pl.w, nc) <= kLdsDefault) {
            T[0] = T[1] = T[2] = cand;
            break;
        }
    }
    return make_bins(T, pl);
}

// Sorts the binned points: keys (bucket * cells + cell) of `total` points with
// the per-bucket counts already on the host (hc) and on the device (count), into
// point ids ordered by (bucket, cell, point index).  Returns the ids buffer.
DevBuf bucket_sort(Ctx &cx, const DevBuf &key, long long total, unsigned cells, const std::vector<unsigned> &hc,
                   const DevBuf &count) {
    const size_t nb = hc.size();
    std::vector<unsigned> ho(nb);
    unsigned long long off = 0;
    for (size_t b = 0; b < nb; ++b) {
        ho[b] = (unsigned)off;
        off += hc[b];
    }
    LAM_REQUIRE((long long)off == total, LAM_ERR_STATE, "bin counts do not add up to the point count");
    DevBuf offset(sizeof(unsigned) * nb), cursor(sizeof(unsigned) * nb);
    DevBuf staged(sizeof(unsigned long long) * std::max<long long>(total, 1));
    DevBuf ids(sizeof(unsigned) * std::max<long long>(total, 1));
    upload_now(cx, offset.ptr, ho.data(), sizeof(unsigned) * nb);
    LAM_HIP(hipMemsetAsync(cursor.ptr, 0, cursor.bytes, cx.stream));
    const unsigned grid = grid_for((size_t)total, 256, 1u << 20);
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(256), 0, cx.stream, key.as<unsigned>(), total, cells,
                       offset.as<unsigned>(), cursor.as<unsigned>(), staged.as<unsigned long long>());
    check_launch();
    constexpr size_t kMaxBlocks = (size_t)1 << 30;
    for (size_t b0 = 0; b0 < nb; b0 += kMaxBlocks) {
        const size_t nblk = std::min(kMaxBlocks, nb - b0);
        hipLaunchKernelGGL(k_bin_order, dim3((unsigned)nblk), dim3(256), 0, cx.stream,
                           staged.as<unsigned long long>(), offset.as<unsigned>(), count.as<unsigned>(),
                           (unsigned)b0, ids.as<unsigned>());
        check_launch();
    }
    LAM_HIP(hipStreamSynchronize(cx.stream));
    return ids;
}

template <class K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024) LAM_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

SortedPoints &get_bins(Ctx &cx, Grid &g, NufftPlan &pl, bool half, const BinParams &bp) {
    std::vector<long long> key = {pl.nf[0], pl.nf[1], pl.nf[2], pl.w, half ? 1 : 0, bp.T[0], bp.T[1], bp.T[2]};
    auto it = g.bins.find(key);
    if (it != g.bins.end()) return *it->second;
    TimedScope ts(cx, "bin_points");
    const long long S = g.R * g.C, H = half ? (S + 1) / 2 : S, total = g.P * H;
    LAM_REQUIRE(total < (1ll << 31), LAM_ERR_ARG, "too many points in one polar grid for 32-bit bins");
    const long long ntiles = (long long)bp.nt[0] * bp.nt[1] * bp.nt[2];
    const long long nkeys = ntiles * bp.T[0] * bp.T[1] * bp.T[2];
    LAM_REQUIRE(nkeys <= (1ll << 32), LAM_ERR_ARG, "fine grid too large for 32-bit bin keys");
    DevBuf key0(sizeof(unsigned) * std::max<long long>(total, 1));
    DevBuf count(sizeof(unsigned) * ntiles);
    LAM_HIP(hipMemsetAsync(count.ptr, 0, count.bytes, cx.stream));
    const PointFrame fr = frame(g, pl);
    const unsigned grid = grid_for((size_t)total, 256, 1u << 20);
    const Geo *geo = g.geo.as<Geo>();
    if (half) {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_bin_keys<W, true>), dim3(grid), dim3(256), 0, cx.stream, geo, g.P,
                                                  (int)g.R, (int)g.C, pl.nf[0], pl.nf[1], pl.nf[2], fr, bp,
                                                  key0.as<unsigned>(), count.as<unsigned>()));
    } else {
        LAM_WIDTH_SWITCH(pl.w, hipLaunchKernelGGL((k_bin_keys<W, false>), dim3(grid), dim3(256), 0, cx.stream, geo,
                                                  g.P, (int)g.R, (int)g.C, pl.nf[0], pl.nf[1], pl.nf[2], fr, bp,
                                                  key0.as<unsigned>(), count.as<unsigned>()));
    }
    check_launch();
    std::vector<unsigned> hc(ntiles);
    LAM_HIP(hipMemcpyAsync(hc.data(), count.ptr, sizeof(unsigned) * ntiles, hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
    // order: tile, then first cell inside the tile, then point index (deterministic)
    auto sp = std::make_unique<SortedPoints>();
    sp->npoints = total;
    sp->ids = bucket_sort(cx, key0, total, (unsigned)(bp.T[0] * bp.T[1] * bp.T[2]), hc, count);
    std::vector<uint4> chunks;
    unsigned long long off = 0;
    for (long long t = 0; t < ntiles; ++t) {
        for (unsigned s0 = 0; s0 < hc[t]; s0 += kChunk)
            chunks.push_back(make_uint4((unsigned)t, (unsigned)(off + s0), std::min(kChunk, hc[t] - s0), 0u));
        off += hc[t];
    }
    sp->nchunks = (long long)chunks.size();
    if (!chunks.empty()) {
        sp->chunks.alloc(sizeof(uint4) * chunks.size());
        LAM_HIP(hipMemcpyAsync(sp->chunks.ptr, chunks.data(), sizeof(uint4) * chunks.size(), hipMemcpyHostToDevice,
                               cx.stream));
        LAM_HIP(hipStreamSynchronize(cx.stream));
    }
    SortedPoints &ref = *sp;
    g.bins.emplace(std::move(key), std::move(sp));
    return ref;
}

// ---- row-factorised path (host)
constexpr int kTileRowsTarget = 16;     // preferred 2-D tile edge of the row path (measured: 16 < 20 < 24)

// Slab memory of one row group: LAMINO_NUFFT_SLAB_BYTES, default 8 GiB, which
// holds the 308 rows x 22 MB of a (720, 144, 720) fine grid with a 615-point
// detector in one group: one y pass that stores each plane once, instead of a
// zero fill and a read-modify-write of the fine grid per group.
size_t slab_budget() {
    const char *e = std::getenv("LAMINO_NUFFT_SLAB_BYTES");
    if (e && *e) {
        const long long v = std::atoll(e);
        if (v > 0) return (size_t)v;
    }
    return (size_t)8 << 30;
}

long long half_rows(const Grid &g) {
    const long long H = (g.R * g.C + 1) / 2;
    return (H - 1) / g.C + 1;
}

// a tile edge that divides n, is a multiple of mult, at least w (the halo reaches one tile
// only), closest to the target; 0 if there is none
int row_tile(int n, int w, int mult, int target) {
    int best = 0;
    for (int d = w; d <= std::min(n, 48); ++d)
        if (n % d == 0 && d % mult == 0 && (best == 0 || std::abs(d - target) < std::abs(best - target)))
            best = d;
    return best;
}

// The 2-D tiling of the row path for tile edges near (tz, tx).
RowGeom row_geom(const NufftPlan &pl, int tz = kTileRowsTarget, int tx = kTileRowsTarget) {
    RowGeom rg;
    const int W = pl.w;
    rg.Tz = row_tile(pl.nf[0], W, 1, tz);
    // the spreader's x block must divide Tx: the largest block that has such a tile
    for (rg.XB = row_xblock_max(W); rg.XB >= 1; --rg.XB)
        if ((rg.Tx = row_tile(pl.nf[2], W, rg.XB, tx)) != 0) break;
    if (rg.XB < 1) rg.XB = 1;
    if (!rg.Tz || !rg.Tx) return rg;
    rg.ntz = pl.nf[0] / rg.Tz;
    rg.ntx = pl.nf[2] / rg.Tx;
    rg.ntiles = (unsigned)(rg.ntz * rg.ntx);
    rg.HBz = rg.Tz + W - 1;
    rg.HBx = rg.Tx + W - 1;
    rg.NH = (W - 1) * rg.Tx + rg.Tz * (W - 1) + (W - 1) * (W - 1);
    rg.fbx = FastDiv((unsigned)rg.HBx);
    rg.fbx3 = FastDiv((unsigned)(3 * rg.HBx));
    return rg;
}

// The three-component projectors take the row path when every projection
// shares the tilt (LAMINO_NUFFT_ROWS=0 turns it off), tile edges dividing the
// fine grid exist and the sort keys fit 32 bits.
bool rows_fit(const Grid &g, const RowGeom &rg) {
    return rg.Tz && half_rows(g) * (long long)rg.ntiles * rg.Tz * rg.Tx <= (1ll << 32);
}

bool use_rows(const Grid &g, const NufftPlan &pl) {
    if (!g.shared_tilt) return false;
    const char *e = std::getenv("LAMINO_NUFFT_ROWS");
    if (e && e[0] == '0') return false;
    return rows_fit(g, row_geom(pl));
}

// Tile edge targets (z, x) the row-path autotuner times; the first is the untuned default.
// (Measured, spread + interpolation of one dataset: at (359, 71, 359) / 615 x 615, 360
// projections, 12 x 15 9.73 ms against 10.11 for the default 16 x 15 and 11.00 for 24 x 20;
// at (179, 43, 179) / 305, 10 x 15 1.36 ms against 1.41.)
constexpr int kRowTileCands[][2] = {{16, 15}, {12, 10}, {12, 15}, {10, 15}, {8, 15}, {10, 20},
                                    {20, 15}, {16, 20}, {20, 20}, {24, 20}};

// The row tiling of this grid's projectors: the tuned one when there is one, else the default.
std::vector<long long> row_tune_key(const Grid &g, const NufftPlan &pl) { return tune_key(g, pl, 0); }

RowGeom row_geom_for(Ctx &cx, const Grid &g, const NufftPlan &pl) {
    std::array<int, 3> T;
    if (lookup_tuned(cx, row_tune_key(g, pl), T)) {
        const RowGeom rg = row_geom(pl, T[0], T[2]);
        if (rg.Tz == T[0] && rg.Tx == T[2] && rows_fit(g, rg)) return rg;
    }
    return row_geom(pl);
}

size_t rows_spread_lds(const RowGeom &rg, int w) {
    const int n = 6 + w + (w + row_xblock_max(w) - 1);
    const size_t sw = (size_t)(n + (6 - n % 4) % 4);   // RowStage<w>::SW
    return (size_t)(kStageRows + 2) * sw * 4 + 3 * (size_t)rg.HBz * rg.HBx * 8;
}
size_t rows_interp_lds(const RowGeom &rg) { return 3 * (size_t)rg.HBz * rg.HBx * 8; }

RowPlan &get_rowplan(Ctx &cx, Grid &g, NufftPlan &pl) {
    const RowGeom rg0 = row_geom_for(cx, g, pl);
    std::vector<long long> key = {pl.nf[0], pl.nf[1], pl.nf[2], pl.w, rg0.Tz, rg0.Tx};
    auto it = g.rowplans.find(key);
    if (it != g.rowplans.end()) return *it->second;
    TimedScope ts(cx, "bin_points");
    const int W = pl.w;
    const long long S = g.R * g.C, H = (S + 1) / 2, total = g.P * H;
    LAM_REQUIRE(total < (1ll << 31), LAM_ERR_ARG, "too many points in one polar grid for 32-bit bins");
    auto rp = std::make_unique<RowPlan>();
    rp->rg = rg0;
    const RowGeom &rg = rp->rg;
    LAM_REQUIRE(rg.Tz > 0, LAM_ERR_STATE, "no row tiling for this fine grid");
    rp->Jn = (int)half_rows(g);
    const long long nkeys = (long long)rp->Jn * rg.ntiles * rg.Tz * rg.Tx;
    LAM_REQUIRE(nkeys <= (1ll << 32), LAM_ERR_ARG, "fine grid too large for 32-bit row bin keys");

    // Row weights from the same coordinate expression as the kernels (ty = (qY vy) sy),
    // in double; iy is the unwrapped first cell, linear in j and so monotonic.
    const PointFrame fr = frame(g, pl);
    const double vy = g.geo_host[0].vy;
    std::vector<float> wyh((size_t)rp->Jn * W);
    std::vector<int> iyh(rp->Jn);
    for (int j = 0; j < rp->Jn; ++j) {
        const double qY = ((double)j - fr.j0) * fr.dY;
        const double ty = (qY * vy) * fr.sy;
        const int iy0 = (int)std::ceil(ty - 0.5 * W);
        iyh[j] = iy0;
        for (int m = 0; m < W; ++m) {
            const double z = ((double)(iy0 + m) - ty) * (2.0 / W);
            const double r = 1.0 - z * z;
            wyh[(size_t)j * W + m] = r > 0 ? (float)std::exp((double)pl.beta * (std::sqrt(r) - 1.0)) : 0.f;
        }
    }
    rp->dir = iyh[rp->Jn - 1] >= iyh[0] ? 1 : -1;
    {
        const int lo = *std::min_element(iyh.begin(), iyh.end()), hi = *std::max_element(iyh.begin(), iyh.end());
        const int len = hi + W - lo;
        if (len < pl.nf[1]) {
            rp->ylo = ((lo % pl.nf[1]) + pl.nf[1]) % pl.nf[1];
            rp->ly = len;
        }
    }
    for (int j = 0; j + 1 < rp->Jn; ++j)
        LAM_REQUIRE(rp->dir * (iyh[j + 1] - iyh[j]) >= 0, LAM_ERR_STATE, "row first cells are not monotonic");
    rp->wy.alloc(sizeof(float) * wyh.size());
    upload_now(cx, rp->wy.ptr, wyh.data(), sizeof(float) * wyh.size());
    rp->iy.alloc(sizeof(int) * iyh.size());
    upload_now(cx, rp->iy.ptr, iyh.data(), sizeof(int) * iyh.size());
    rp->iy_host = iyh;

    const long long ncount = (long long)rp->Jn * rg.ntiles;
    DevBuf key0(sizeof(unsigned) * std::max<long long>(total, 1));
    rp->count.alloc(sizeof(unsigned) * ncount);
    LAM_HIP(hipMemsetAsync(rp->count.ptr, 0, rp->count.bytes, cx.stream));
    const unsigned grid = grid_for((size_t)total, 256, 1u << 20);
    LAM_WIDTH_SWITCH(W, hipLaunchKernelGGL((k_bin_keys_rows<W>), dim3(grid), dim3(256), 0, cx.stream,
                                           g.geo.as<Geo>(), g.P, (int)g.R, (int)g.C, pl.nf[0], pl.nf[2], fr, rg,
                                           key0.as<unsigned>(), rp->count.as<unsigned>()));
    check_launch();
    std::vector<unsigned> hc(ncount);
    LAM_HIP(hipMemcpyAsync(hc.data(), rp->count.ptr, sizeof(unsigned) * ncount, hipMemcpyDeviceToHost, cx.stream));
    LAM_HIP(hipStreamSynchronize(cx.stream));
    // order: row, tile, first (z, x) cell inside the tile, then point index (deterministic)
    rp->ids = bucket_sort(cx, key0, total, (unsigned)(rg.Tz * rg.Tx), hc, rp->count);
    // one chunk per non-empty (row, tile): its workgroup alone writes that tile's
    // interior and halo record
    std::vector<uint4> chunks;
    rp->row_chunk.assign((size_t)rp->Jn + 1, 0);
    unsigned long long off = 0;
    for (long long jt = 0; jt < ncount; ++jt) {
        if (jt % rg.ntiles == 0) rp->row_chunk[jt / rg.ntiles] = (long long)chunks.size();
        if (hc[jt]) chunks.push_back(make_uint4((unsigned)jt, (unsigned)off, hc[jt], 0u));
        off += hc[jt];
    }
    rp->nchunks = (long long)chunks.size();
    rp->row_chunk[rp->Jn] = rp->nchunks;
    // bit table of the non-empty (row, tile) pairs, tile-major: the y passes test it
    // per fine-grid cell and row
    rp->nzw = (rp->Jn + 31) / 32;
    {
        std::vector<unsigned> bits((size_t)rg.ntiles * rp->nzw, 0u);
        for (long long jt = 0; jt < ncount; ++jt)
            if (hc[jt]) {
                const long long j = jt / rg.ntiles, t = jt - j * rg.ntiles;
                bits[(size_t)t * rp->nzw + (size_t)(j >> 5)] |= 1u << (j & 31);
            }
        rp->nz.alloc(sizeof(unsigned) * bits.size());
        upload_now(cx, rp->nz.ptr, bits.data(), sizeof(unsigned) * bits.size());
    }
    if (!chunks.empty()) {
        rp->chunks.alloc(sizeof(uint4) * chunks.size());
        LAM_HIP(hipMemcpyAsync(rp->chunks.ptr, chunks.data(), sizeof(uint4) * chunks.size(), hipMemcpyHostToDevice,
                               cx.stream));
        LAM_HIP(hipStreamSynchronize(cx.stream));
    }
    RowPlan &ref = *rp;
    g.rowplans.emplace(std::move(key), std::move(rp));
    return ref;
}

size_t row_slab_bytes(const NufftPlan &pl) { return (size_t)pl.nf[0] * pl.nf[2] * 3 * sizeof(float2); }
size_t row_halo_bytes(const RowGeom &rg) { return (size_t)rg.ntiles * rg.NH * 3 * sizeof(float2); }

// rows per group: as many slabs (and halo records) as the budget holds
int rows_per_group(const RowPlan &rp, const NufftPlan &pl) {
    const size_t per_row = row_slab_bytes(pl) + row_halo_bytes(rp.rg);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)rp.Jn, slab_budget() / per_row));
}

}  // namespace

// Upper bound of what the row path allocates for one grid (see nufft.h).
RowBytes row_path_bytes(const long long dims[3], double eps, long long P, long long R, long long C) {
    NufftPlan pl;   // shapes only
    pl.w = width_for_eps(eps);
    for (int a = 0; a < 3; ++a) pl.nf[a] = (int)nufft_fine_len(dims[a], eps);
    RowBytes b;
    if (!row_geom(pl).Tz) return b;
    const long long H = (R * C + 1) / 2, Jn = (H - 1) / C + 1, half = P * H;
    // the largest over the tilings the autotuner may choose
    for (const auto &t : kRowTileCands) {
        const RowGeom rg = row_geom(pl, t[0], t[1]);
        if (!rg.Tz) continue;
        const size_t per_row = row_slab_bytes(pl) + row_halo_bytes(rg);
        const size_t groups_rows = std::max<size_t>(1, std::min<size_t>((size_t)Jn, slab_budget() / per_row));
        b.slab = std::max(b.slab, groups_rows * per_row);
        const size_t cells = (size_t)Jn * rg.ntiles;
        // ids, counts + chunks, row weights, the non-empty bit table
        b.kept = std::max(b.kept, 4 * (size_t)half + 20 * cells + 8 * (size_t)Jn * pl.w + cells / 8 + 4 * rg.ntiles);
    }
    b.setup = 16 * (size_t)half + ((size_t)16 << 20);                // sort keys / spare ids, sort scratch
    return b;
}

namespace {

// Columns per workgroup of the windowed contraction for a group of rows (0: the
// window does not fit; use the row walk), with the window's first plane and
// length.  The window [span][columns] of float2 stays within 64 KiB of LDS.
struct YWinLaunch {
    int cols = 0, lo = 0, span = 0;
};
YWinLaunch ywin_launch(const RowPlan &rp, int W, int j0, int nrows) {
    YWinLaunch y;
    const int first = rp.dir > 0 ? j0 : j0 + nrows - 1, last = rp.dir > 0 ? j0 + nrows - 1 : j0;
    y.lo = rp.iy_host[first];
    y.span = rp.iy_host[last] + W - y.lo;
    for (int c : {256, 128, 64})
        if ((size_t)y.span * c * sizeof(float2) <= 65536) { y.cols = c; break; }
    return y;
}

template <int W>
void interp_rows_w(Ctx &cx, Grid &g, NufftPlan &pl, float2 *c, RowPlan &rp) {
    const int Jg = rows_per_group(rp, pl);
    pl.slab.ensure(row_slab_bytes(pl) * Jg);
    const PointFrame fr = frame(g, pl);
    const size_t lds = rows_interp_lds(rp.rg);
    allow_lds(&k_rows_interp<W>, lds);
    const long long Hh = (g.R * g.C + 1) / 2;
    const FastDiv fH((unsigned)Hh), fC((unsigned)g.C);
    allow_lds(&k_rows_contract_win<W>, 65536);
    const unsigned cgrid = grid_for((size_t)pl.nf[0] * pl.nf[2] * 3, 256, 1u << 20);
    for (int j0 = 0; j0 < rp.Jn; j0 += Jg) {
        const int nrows = std::min(Jg, rp.Jn - j0);
        {
            TimedScope ts(cx, "y_contract");
            const YWinLaunch yw = ywin_launch(rp, W, j0, nrows);
            if (yw.cols) {
                const size_t cols_total = (size_t)pl.nf[0] * pl.nf[2] * 3;
                hipLaunchKernelGGL((k_rows_contract_win<W>), dim3((unsigned)((cols_total + yw.cols - 1) / yw.cols)),
                                   dim3(yw.cols), (size_t)yw.span * yw.cols * sizeof(float2), cx.stream,
                                   pl.fine.as<float2>(), pl.slab.as<float2>(), rp.wy.as<float>(), rp.iy.as<int>(),
                                   pl.nf[0], pl.nf[1], pl.nf[2], j0, nrows, rp.dir, yw.lo, yw.span);
            } else {
                hipLaunchKernelGGL((k_rows_contract<W>), dim3(cgrid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                                   pl.slab.as<float2>(), rp.wy.as<float>(), rp.iy.as<int>(), pl.nf[0], pl.nf[1],
                                   pl.nf[2], j0, nrows, rp.dir);
            }
            check_launch();
        }
        const long long c0 = rp.row_chunk[j0], c1 = rp.row_chunk[j0 + nrows];
        if (c1 > c0) {
            TimedScope ts(cx, "interp_rows2d");
            hipLaunchKernelGGL((k_rows_interp<W>), dim3((unsigned)(c1 - c0)), dim3(kRowsInterpThreads), lds, cx.stream,
                               pl.slab.as<float2>(), c, g.geo.as<Geo>(), g.P, (int)g.R, (int)g.C, pl.nf[0],
                               pl.nf[2], pl.beta, fr, rp.rg, j0, fH, fC, rp.ids.as<unsigned>(),
                               rp.chunks.as<uint4>() + c0);
            check_launch();
        }
    }
}

template <int W>
void spread_rows_w(Ctx &cx, Grid &g, NufftPlan &pl, const float2 *c, RowPlan &rp, bool paired, bool stores) {
    const int Jg = rows_per_group(rp, pl);
    pl.slab.ensure(row_slab_bytes(pl) * Jg);
    pl.halo.ensure(row_halo_bytes(rp.rg) * Jg);
    const PointFrame fr = frame(g, pl);
    const size_t lds = rows_spread_lds(rp.rg, W);
    allow_lds(&k_rows_spread<W>, lds);
    // point ids are below 2^31 (the row plan requires it), so 32-bit multiply-shift
    // division splits them into (projection, index) and (row, column)
    const long long Hh = (g.R * g.C + 1) / 2;
    const FastDiv fH((unsigned)Hh), fC((unsigned)g.C);
    const unsigned cgrid = grid_for((size_t)pl.nf[0] * pl.nf[2] * 3, 256, 1u << 20);
    for (int j0 = 0; j0 < rp.Jn; j0 += Jg) {
        const int nrows = std::min(Jg, rp.Jn - j0);
        const long long c0 = rp.row_chunk[j0], c1 = rp.row_chunk[j0 + nrows];
        if (c1 > c0) {
            TimedScope ts(cx, "spread_rows2d");
            const unsigned nch = (unsigned)(c1 - c0);
            hipLaunchKernelGGL((k_rows_spread<W>), dim3((nch + kSpreadChunks - 1) / kSpreadChunks), dim3(64), lds,
                               cx.stream, pl.slab.as<float2>(), pl.halo.as<float2>(), c, g.geo.as<Geo>(), (int)g.R,
                               (int)g.C, pl.nf[0], pl.nf[2], pl.beta, fr, rp.rg, j0, fH, fC, rp.ids.as<unsigned>(),
                               rp.chunks.as<uint4>() + c0, nch, paired ? 1 : 0);
            check_launch();
        }
        TimedScope ts(cx, "y_expand");
        hipLaunchKernelGGL((k_rows_expand<W>), dim3(cgrid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           pl.slab.as<float2>(), pl.halo.as<float2>(), rp.nz.as<unsigned>(), rp.nzw, rp.wy.as<float>(),
                           rp.iy.as<int>(), pl.nf[0], pl.nf[1], pl.nf[2], rp.rg, j0, nrows, rp.dir, stores ? 0 : 1);
        check_launch();
    }
}

template <int W, bool HALF>
void spread_w(Ctx &cx, Grid &g, NufftPlan &pl, const float2 *c, const BinParams &bp, SortedPoints &sp, bool paired) {
    constexpr int NC = HALF ? 3 : 1;
    const size_t lds = spread_lds_bytes(bp.T, W, NC);
    allow_lds(&k_spread_tiled<W, HALF>, lds);
    hipLaunchKernelGGL((k_spread_tiled<W, HALF>), dim3((unsigned)sp.nchunks), dim3(256), lds, cx.stream,
                       pl.fine.as<float2>(), c, g.geo.as<Geo>(), g.P, (int)g.R, (int)g.C, pl.nf[0], pl.nf[1], pl.nf[2],
                       pl.beta, frame(g, pl), bp, sp.ids.as<unsigned>(), sp.chunks.as<uint4>(), paired ? 1 : 0);
}

template <int W, bool HALF>
void interp_w(Ctx &cx, Grid &g, NufftPlan &pl, float2 *c, const BinParams &bp, SortedPoints &sp) {
    constexpr int NC = HALF ? 3 : 1;
    const size_t lds = interp_lds_bytes(bp.T, W, NC);
    allow_lds(&k_interp_tiled<W, HALF>, lds);
    hipLaunchKernelGGL((k_interp_tiled<W, HALF>), dim3((unsigned)sp.nchunks), dim3(256), lds, cx.stream,
                       pl.fine.as<float2>(), c, g.geo.as<Geo>(), g.P, (int)g.R, (int)g.C, pl.nf[0], pl.nf[1], pl.nf[2],
                       pl.beta, frame(g, pl), bp, sp.ids.as<unsigned>(), sp.chunks.as<uint4>());
}

// paired: c holds detector-pair spectra (det_fft2_real_pairs), three-component form only
// add: accumulate into the fine grid's current contents (no zero fill, no stores)
void launch_spread(Ctx &cx, Grid &g, NufftPlan &pl, const float2 *c, bool half, bool paired = false,
                   bool add = false) {
    const int nc = half ? 3 : 1;
    if (half && use_rows(g, pl)) {
        RowPlan &rp = get_rowplan(cx, g, pl);
        // One row group whose y window does not wrap writes every plane of the window
        // once, and the transforms that follow read only the window: no zeroing.
        const bool stores = !add && rows_per_group(rp, pl) >= rp.Jn && rp.ly > 0;
        if (!stores && !add)
            LAM_HIP(hipMemsetAsync(pl.fine.ptr, 0, sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * 3,
                                   cx.stream));
        TimedScope ts(cx, "spread");
        LAM_WIDTH_SWITCH(pl.w, (spread_rows_w<W>(cx, g, pl, c, rp, paired, stores)));
        return;
    }
    const BinParams bp = bin_params(cx, g, pl, nc);
    SortedPoints &sp = get_bins(cx, g, pl, half, bp);
    if (!add)
        LAM_HIP(hipMemsetAsync(pl.fine.ptr, 0, sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * nc,
                               cx.stream));
    if (sp.nchunks == 0) return;
    TimedScope ts(cx, "spread");
    if (half) {
        LAM_WIDTH_SWITCH(pl.w, (spread_w<W, true>(cx, g, pl, c, bp, sp, paired)));
    } else {
        LAM_WIDTH_SWITCH(pl.w, (spread_w<W, false>(cx, g, pl, c, bp, sp, false)));
    }
    check_launch();
}

void launch_interp(Ctx &cx, Grid &g, NufftPlan &pl, float2 *c, bool half) {
    const int nc = half ? 3 : 1;
    if (half && use_rows(g, pl)) {
        RowPlan &rp = get_rowplan(cx, g, pl);
        TimedScope ts(cx, "interp");
        LAM_WIDTH_SWITCH(pl.w, (interp_rows_w<W>(cx, g, pl, c, rp)));
        return;
    }
    const BinParams bp = bin_params(cx, g, pl, nc);
    SortedPoints &sp = get_bins(cx, g, pl, half, bp);
    if (sp.nchunks == 0) return;
    TimedScope ts(cx, "interp");
    if (half) {
        LAM_WIDTH_SWITCH(pl.w, (interp_w<W, true>(cx, g, pl, c, bp, sp)));
    } else {
        LAM_WIDTH_SWITCH(pl.w, (interp_w<W, false>(cx, g, pl, c, bp, sp)));
    }
    check_launch();
}

void ensure_fine(NufftPlan &pl, int nc) {
    pl.fine.ensure(sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * nc);
}

void deapod_pad(Ctx &cx, NufftPlan &pl, const float *const m[3], const float2 *cin, int nc) {
    TimedScope ts(cx, "deapod_pad");
    const long long total = pl.n[0] * pl.n[1] * pl.n[2];
    const unsigned grid = grid_for((size_t)total, 256, 1u << 16);
    if (nc == 3)
        hipLaunchKernelGGL((k_deapod_pad<3, true>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(), m[0],
                           m[1], m[2], nullptr, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0], pl.nf[1],
                           pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(), pl.corr[2].as<float>());
    else
        hipLaunchKernelGGL((k_deapod_pad<1, false>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           nullptr, nullptr, nullptr, cin, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>());
    check_launch();
}

void crop_deapod(Ctx &cx, NufftPlan &pl, float *const out[3], float2 *cout, int nc, float scale, bool accumulate) {
    TimedScope ts(cx, "crop_deapod");
    const long long total = pl.n[0] * pl.n[1] * pl.n[2];
    const unsigned grid = grid_for((size_t)total, 256, 1u << 16);
    if (nc == 3)
        hipLaunchKernelGGL((k_crop_deapod<3, true>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           out[0], out[1], out[2], nullptr, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>(), scale, accumulate ? 1 : 0);
    else
        hipLaunchKernelGGL((k_crop_deapod<1, false>), dim3(grid), dim3(256), 0, cx.stream, pl.fine.as<float2>(),
                           nullptr, nullptr, nullptr, cout, (int)pl.n[0], (int)pl.n[1], (int)pl.n[2], pl.nf[0],
                           pl.nf[1], pl.nf[2], pl.corr[0].as<float>(), pl.corr[1].as<float>(),
                           pl.corr[2].as<float>(), scale, 0);
    check_launch();
}

// Fine-grid 3-D FFT restricted to what the NUFFT needs.  Along an axis with n
// modes the mode region of the fine grid is the wrapped range
// [0, n - n/2) U [nf - n/2, nf).
//  forward (type 2): the input is zero outside the mode region, so the x pass
//  runs only on (z, y) mode lines and every pass reads only mode entries;
//  inverse (type 1): only the mode region of the output is kept, so later
//  passes run on fewer lines and every pass stores only mode entries.
// The fine-grid y planes the three-component projectors of this grid use: the
// rows' window on the row path, all planes otherwise.
struct YWin {
    int lo = 0, len = 0;
};
YWin y_window(Ctx &cx, Grid &g, NufftPlan &pl) {
    YWin w;
    if (!use_rows(g, pl)) return w;
    const RowPlan &rp = get_rowplan(cx, g, pl);
    w.lo = rp.ylo;
    w.len = rp.ly;
    return w;
}

// 3-D transform of the fine grid between the mode region and the full grid.
// [ylo, ylo + ly) mod nfy (ly > 0): the only y planes the caller reads after
// the forward transform / the only nonzero ones before the inverse (the rows
// of a row-factorised projector); the z pass then runs over those planes only.
void fine_fft(Ctx &cx, NufftPlan &pl, int nc, int sign, int ylo = 0, int ly = 0) {
    const int nfz = pl.nf[0], nfy = pl.nf[1], nfx = pl.nf[2];
    const int nz = (int)pl.n[0], ny = (int)pl.n[1], nx = (int)pl.n[2];
    const int hz = nz - nz / 2, hy = ny - ny / 2, hx = nx - nx / 2;   // head counts
    float2 *f = pl.fine.as<float2>();
    const long long sX = nc, sY = (long long)nfx * nc, sZ = (long long)nfy * nfx * nc;
    auto x_pass = [&](bool fwd) {
        AxisPass p;
        p.in = p.out = f;
        p.is1 = p.os1 = sZ; p.is2 = p.os2 = sY; p.isi = p.osi = 1; p.isn = p.osn = sX;
        p.c1 = nz; p.o1_split = hz; p.o1_gap = nfz - nz;
        p.c2 = ny; p.o2_split = hy; p.o2_gap = nfy - ny;
        p.ci = nc;
        if (fwd) { p.nv = hx; p.nv2 = nx / 2; } else { p.ns = hx; p.ns2 = nx / 2; }
        run_axis(cx, get_fft_plan(cx, nfx, sign), p);
    };
    auto y_pass = [&](bool fwd) {
        AxisPass p;
        p.in = p.out = f;
        p.is1 = p.os1 = sZ; p.isi = p.osi = 1; p.isn = p.osn = sY;
        p.c1 = nz; p.o1_split = hz; p.o1_gap = nfz - nz;
        p.c2 = 1; p.ci = nfx * nc;
        if (fwd) { p.nv = hy; p.nv2 = ny / 2; } else { p.ns = hy; p.ns2 = ny / 2; }
        if (ly > 0) {
            if (fwd) { p.st_lo = ylo; p.st_len = ly; }
            else { p.ld_lo = ylo; p.ld_len = ly; }
        }
        run_axis(cx, get_fft_plan(cx, nfy, sign), p);
    };
    auto z_pass = [&](bool fwd) {
        AxisPass p;
        p.in = p.out = f;
        p.is2 = p.os2 = sY; p.isi = p.osi = 1; p.isn = p.osn = sZ;
        p.c1 = 1; p.c2 = nfy; p.ci = nfx * nc;
        if (ly > 0) {
            // lines (y, x) for y in the window; line offsets must stay non-negative (the
            // pass marks lines outside the set with -1), so a window wrapping past nfy
            // is taken as [0, e2) followed by [ylo, nfy)
            p.c2 = ly;
            if (ylo + ly <= nfy) {
                p.in = p.out = f + (long long)ylo * sY;
            } else {
                const int e2 = ylo + ly - nfy;
                p.o2_split = e2;
                p.o2_gap = ylo - e2;
            }
        }
        if (fwd) { p.nv = hz; p.nv2 = nz / 2; } else { p.ns = hz; p.ns2 = nz / 2; }
        run_axis(cx, get_fft_plan(cx, nfz, sign), p);
    };
    if (sign < 0) { x_pass(true); y_pass(true); z_pass(true); }
    else { z_pass(false); y_pass(false); x_pass(false); }
}

void check_dims(const long long dims[3]) {
    for (int a = 0; a < 3; ++a)
        LAM_REQUIRE(dims[a] >= 1 && dims[a] <= (1 << 20), LAM_ERR_ARG, "invalid volume dimensions");
}

}  // namespace

Grid::~Grid() = default;

int width_for_eps(double eps) {
    if (!(eps > 0)) eps = 1e-5;
    int w = (int)std::ceil(-std::log10(eps / 10.0));
    if (w % 2) ++w;
    return std::max(4, std::min(12, w));
}

size_t nufft_fine_len(long long n, double eps) {
    const size_t w = (size_t)width_for_eps(eps);
    return next_smooth_even(std::max<size_t>(2 * (size_t)n, 2 * w));
}

NufftPlan &get_nufft_plan(Ctx &cx, const long long dims[3], double eps) {
    check_dims(dims);
    const int w = width_for_eps(eps);
    auto key = std::make_tuple(dims[0], dims[1], dims[2] * 64 + w, 0);
    auto it = cx.nufft_plans.find(key);
    if (it != cx.nufft_plans.end()) return *it->second;
    auto pl = std::make_unique<NufftPlan>();
    pl->w = w;
    pl->beta = (float)(2.30 * w);
    const int nq = 200;
    std::vector<double> xq(nq), wq(nq);
    gauss_legendre(nq, xq.data(), wq.data());
    const double pi = 3.14159265358979323846;
    for (int a = 0; a < 3; ++a) {
        pl->n[a] = dims[a];
        const size_t f = nufft_fine_len(dims[a], eps);
        pl->nf[a] = (int)f;
        std::vector<float> corr(dims[a]);
        for (long long i = 0; i < dims[a]; ++i) {
            const long long k = i - dims[a] / 2;
            double s = 0;
            for (int q = 0; q < nq; ++q) {
                const double z = xq[q];
                const double phi = std::exp((double)pl->beta * (std::sqrt(1.0 - z * z) - 1.0));
                s += wq[q] * phi * std::cos(pi * (double)k * w * z / (double)f);
            }
            corr[i] = (float)(1.0 / (0.5 * w * s));
        }
        pl->corr[a].alloc(sizeof(float) * dims[a]);
        upload_now(cx, pl->corr[a].ptr, corr.data(), sizeof(float) * dims[a]);
    }
    NufftPlan &ref = *pl;
    cx.nufft_plans.emplace(key, std::move(pl));
    return ref;
}

std::unique_ptr<Grid> make_grid(Ctx &cx, const double *alpha, const double *gamma, long long P, long long R,
                                long long C) {
    LAM_REQUIRE(P >= 1 && R >= 1 && C >= 1, LAM_ERR_ARG, "invalid polar grid shape");
    LAM_REQUIRE(R * C < (1ll << 40), LAM_ERR_ARG, "polar grid too large");
    auto g = std::make_unique<Grid>();
    g->cx = &cx;
    g->P = P; g->R = R; g->C = C;
    g->alpha.assign(alpha, alpha + P);
    g->gamma.assign(gamma, gamma + P);
    g->shared_tilt = std::all_of(alpha, alpha + P, [&](double a) { return a == alpha[0]; });
    g->geo_host.resize(P);
    for (long long p = 0; p < P; ++p) {
        const double a = alpha[p], gm = gamma[p];
        Geo e;
        e.ux = std::cos(gm);
        e.uz = -std::sin(gm);
        e.vx = std::cos(a) * std::sin(gm);
        e.vy = -std::sin(a);
        e.vz = std::cos(a) * std::cos(gm);
        e.d0 = std::sin(gm) * std::sin(a);
        e.d1 = std::cos(a);
        e.d2 = std::cos(gm) * std::sin(a);
        g->geo_host[p] = e;
    }
    g->geo.alloc(sizeof(Geo) * P);
    upload_now(cx, g->geo.ptr, g->geo_host.data(), sizeof(Geo) * P);
    return g;
}

void nufft2_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, float2 *c, const float2 *f) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 1);
    deapod_pad(cx, pl, nullptr, f, 1);
    fine_fft(cx, pl, 1, -1);
    launch_interp(cx, g, pl, c, false);
}

void nufft1_complex(Ctx &cx, Grid &g, const long long dims[3], double eps, const float2 *c, float2 *f) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 1);
    launch_spread(cx, g, pl, c, false);
    fine_fft(cx, pl, 1, +1);
    crop_deapod(cx, pl, nullptr, f, 1, 1.0f, false);
}

// Real projections go through the detector transform two planes at a time
// when the mirrored detector point is the transform's negative frequency.
bool paired_planes(const Grid &g) { return (g.R & 1) && (g.C & 1) && g.P > 1; }

void forward(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const m[3], float *proj) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    deapod_pad(cx, pl, m, nullptr, 3);
    const YWin yw = y_window(cx, g, pl);
    fine_fft(cx, pl, 3, -1, yw.lo, yw.len);
    launch_interp(cx, g, pl, cbuf, true);
    const float scale = 1.0f / (float)(g.R * g.C);
    if (paired_planes(g)) det_ifft2_real_pairs(cx, cbuf, proj, (int)g.P, (int)g.R, (int)g.C, +1, scale);
    else det_fft2(cx, cbuf, false, proj, true, (int)g.P, (int)g.R, (int)g.C, +1, scale, cbuf);
}

void adjoint(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *proj, float *const m[3],
             bool accumulate) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    const bool paired = paired_planes(g);
    if (paired) det_fft2_real_pairs(cx, proj, cbuf, (int)g.P, (int)g.R, (int)g.C, -1);
    else det_fft2(cx, proj, true, cbuf, false, (int)g.P, (int)g.R, (int)g.C, -1, 1.0f, cbuf);
    launch_spread(cx, g, pl, cbuf, true, paired);
    const YWin yw = y_window(cx, g, pl);
    fine_fft(cx, pl, 3, +1, yw.lo, yw.len);
    crop_deapod(cx, pl, m, nullptr, 3, 1.0f / (float)(g.R * g.C), accumulate);
}

namespace {
// The y window shared by several grids' projectors: theirs when all agree, else all planes.
YWin common_window(Ctx &cx, Grid *const *grids, int ngrids, NufftPlan &pl) {
    YWin w = y_window(cx, *grids[0], pl);
    for (int k = 1; k < ngrids; ++k) {
        const YWin v = y_window(cx, *grids[k], pl);
        if (v.lo != w.lo || v.len != w.len) return YWin{};
    }
    return w;
}
}  // namespace

void forward_multi(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], double eps,
                   const float *const m[3], float *const *proj) {
    LAM_REQUIRE(ngrids >= 1, LAM_ERR_ARG, "no grids");
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    size_t M = 0;
    for (int k = 0; k < ngrids; ++k) M = std::max(M, (size_t)(grids[k]->P * grids[k]->R * grids[k]->C));
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    deapod_pad(cx, pl, m, nullptr, 3);
    const YWin yw = common_window(cx, grids, ngrids, pl);
    fine_fft(cx, pl, 3, -1, yw.lo, yw.len);
    for (int k = 0; k < ngrids; ++k) {
        Grid &g = *grids[k];
        launch_interp(cx, g, pl, cbuf, true);
        const float scale = 1.0f / (float)(g.R * g.C);
        if (paired_planes(g)) det_ifft2_real_pairs(cx, cbuf, proj[k], (int)g.P, (int)g.R, (int)g.C, +1, scale);
        else det_fft2(cx, cbuf, false, proj[k], true, (int)g.P, (int)g.R, (int)g.C, +1, scale, cbuf);
    }
}

void adjoint_multi(Ctx &cx, Grid *const *grids, int ngrids, const long long dims[3], double eps,
                   const float *const *proj, float *const m[3], bool accumulate) {
    LAM_REQUIRE(ngrids >= 1, LAM_ERR_ARG, "no grids");
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    size_t M = 0;
    for (int k = 0; k < ngrids; ++k) M = std::max(M, (size_t)(grids[k]->P * grids[k]->R * grids[k]->C));
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    const YWin yw = common_window(cx, grids, ngrids, pl);
    // the first grid stores its window (or zero-fills); with different windows the
    // union is not what the first grid writes, so the whole fine grid starts at zero
    bool add = false;
    if (ngrids > 1 && yw.len == 0) {
        LAM_HIP(hipMemsetAsync(pl.fine.ptr, 0, sizeof(float2) * (size_t)pl.nf[0] * pl.nf[1] * pl.nf[2] * 3,
                               cx.stream));
        add = true;
    }
    const float scale = 1.0f / (float)(grids[0]->R * grids[0]->C);
    for (int k = 0; k < ngrids; ++k) {
        Grid &g = *grids[k];
        LAM_REQUIRE(g.R * g.C == grids[0]->R * grids[0]->C, LAM_ERR_SHAPE,
                    "summed adjoints need one detector size (the 1/(R C) scale is shared)");
        const bool paired = paired_planes(g);
        if (paired) det_fft2_real_pairs(cx, proj[k], cbuf, (int)g.P, (int)g.R, (int)g.C, -1);
        else det_fft2(cx, proj[k], true, cbuf, false, (int)g.P, (int)g.R, (int)g.C, -1, 1.0f, cbuf);
        launch_spread(cx, g, pl, cbuf, true, paired, add);
        add = true;
    }
    fine_fft(cx, pl, 3, +1, yw.lo, yw.len);
    crop_deapod(cx, pl, m, nullptr, 3, scale, accumulate);
}

void sysmat_nufft(Ctx &cx, Grid &g, const long long dims[3], double eps, const float *const x[3],
                  float *const y[3]) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    deapod_pad(cx, pl, x, nullptr, 3);
    const YWin yw = y_window(cx, g, pl);
    fine_fft(cx, pl, 3, -1, yw.lo, yw.len);
    launch_interp(cx, g, pl, cbuf, true);
    launch_spread(cx, g, pl, cbuf, true);
    fine_fft(cx, pl, 3, +1, yw.lo, yw.len);
    crop_deapod(cx, pl, y, nullptr, 3, 1.0f / (float)(g.R * g.C), false);
}

// Times the three-component spreader and interpolator of this grid on the
// fine grid of `dims` for a list of tile shapes (cubic, and flattened ones for
// thin or long axes) that fit the LDS budget, and records the fastest in the
// context's table, where every later spread/interp of the same shapes finds
// it.  Results do not depend on the tile shape beyond FP32 summation order.
bool bins_tuned(Ctx &cx, Grid &g, const long long dims[3], double eps) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    std::array<int, 3> T;
    return lookup_tuned(cx, tune_key(g, pl, 3), T);
}

std::array<int, 3> autotune_bins(Ctx &cx, Grid &g, const long long dims[3], double eps, double *best_ms) {
    NufftPlan &pl = get_nufft_plan(cx, dims, eps);
    if (use_rows(g, pl)) {
        // The projectors of this grid take the row path: time one spread +
        // interpolation for each distinct 2-D tiling the candidate edges give (each
        // with its own point bins), keep the fastest's bins and record its tiling
        // as (Tz, W, Tx) in the row table; the 3-D table gets the default 3-D tile
        // so that the grid counts as tuned.
        ensure_fine(pl, 3);
        const size_t M = (size_t)(g.P * g.R * g.C);
        cx.work[0].ensure(sizeof(float2) * M);
        float2 *cbuf = cx.work[0].as<float2>();
        LAM_HIP(hipMemsetAsync(cbuf, 0, sizeof(float2) * M, cx.stream));
        const auto rkey = row_tune_key(g, pl);
        hipEvent_t e0, e1;
        LAM_HIP(hipEventCreate(&e0));
        LAM_HIP(hipEventCreate(&e1));
        const bool timing = cx.timer.on;
        cx.timer.on = false;
        const bool verbose = std::getenv("LAMINO_TUNE_VERBOSE") != nullptr;
        std::array<int, 3> best = {0, 0, 0};
        float best_t = 0.f;
        std::vector<std::pair<int, int>> seen;
        auto plan_key = [&](int tz, int tx) {
            return std::vector<long long>{pl.nf[0], pl.nf[1], pl.nf[2], pl.w, tz, tx};
        };
        try {
            for (const auto &cand : kRowTileCands) {
                const RowGeom rg = row_geom(pl, cand[0], cand[1]);
                if (!rows_fit(g, rg)) continue;
                if (std::find(seen.begin(), seen.end(), std::make_pair(rg.Tz, rg.Tx)) != seen.end()) continue;
                seen.emplace_back(rg.Tz, rg.Tx);
                cx.bin_tune[rkey] = {rg.Tz, pl.w, rg.Tx};
                launch_spread(cx, g, pl, cbuf, true);   // builds this tiling's bins
                float t = 0.f;
                for (int rep = 0; rep < 2; ++rep) {
                    LAM_HIP(hipEventRecord(e0, cx.stream));
                    launch_spread(cx, g, pl, cbuf, true);
                    launch_interp(cx, g, pl, cbuf, true);
                    LAM_HIP(hipEventRecord(e1, cx.stream));
                    LAM_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    LAM_HIP(hipEventElapsedTime(&ms, e0, e1));
                    t = rep == 0 ? ms : std::min(t, ms);
                }
                if (verbose) std::fprintf(stderr, "row tiling %d x %d: %.3f ms\n", rg.Tz, rg.Tx, t);
                if (best[0] == 0 || t < best_t) {
                    if (best[0]) g.rowplans.erase(plan_key(best[0], best[2]));
                    best = {rg.Tz, pl.w, rg.Tx};
                    best_t = t;
                } else {
                    g.rowplans.erase(plan_key(rg.Tz, rg.Tx));
                }
            }
        } catch (...) {
            cx.timer.on = timing;
            cx.bin_tune.erase(rkey);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        cx.timer.on = timing;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        LAM_REQUIRE(best[0] != 0, LAM_ERR_STATE, "no row tiling fits this grid");
        const BinParams bp = bin_params(cx, g, pl, 3);
        const std::array<int, 3> T3 = {bp.T[0], bp.T[1], bp.T[2]};
        cx.bin_tune[rkey] = best;
        cx.bin_tune[tune_key(g, pl, 3)] = T3;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[rkey] = best;
            g_tune[tune_key(g, pl, 3)] = T3;
        }
        if (best_ms) *best_ms = best_t;
        return best;
    }
    ensure_fine(pl, 3);
    const size_t M = (size_t)(g.P * g.R * g.C);
    cx.work[0].ensure(sizeof(float2) * M);
    float2 *cbuf = cx.work[0].as<float2>();
    LAM_HIP(hipMemsetAsync(cbuf, 0, sizeof(float2) * M, cx.stream));
    const auto key = tune_key(g, pl, 3);
    static const int kCands[][3] = {{5, 5, 5}, {6, 6, 6}, {7, 7, 7}, {8, 8, 8}, {6, 4, 8},
                                    {4, 6, 8}, {6, 6, 8}, {4, 8, 8}, {8, 4, 6}, {5, 6, 7}};
    hipEvent_t e0, e1;
    LAM_HIP(hipEventCreate(&e0));
    LAM_HIP(hipEventCreate(&e1));
    std::array<int, 3> best = {0, 0, 0};
    float best_t = 0.f;
    const bool timing = cx.timer.on;
    cx.timer.on = false;
    try {
        for (const auto &T : kCands) {
            if (!bins_fit(T, pl, 3)) continue;
            cx.bin_tune[key] = {T[0], T[1], T[2]};
            launch_spread(cx, g, pl, cbuf, true);   // builds this shape's bins
            float t = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                LAM_HIP(hipEventRecord(e0, cx.stream));
                launch_spread(cx, g, pl, cbuf, true);
                launch_interp(cx, g, pl, cbuf, true);
                LAM_HIP(hipEventRecord(e1, cx.stream));
                LAM_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                LAM_HIP(hipEventElapsedTime(&ms, e0, e1));
                t = rep == 0 ? ms : std::min(t, ms);
            }
            // keep at most two shapes' bins alive: the best so far and this one
            const std::array<int, 3> cur = {T[0], T[1], T[2]};
            std::array<int, 3> drop = cur;
            if (best[0] == 0 || t < best_t) {
                drop = best;
                best = cur;
                best_t = t;
            }
            if (drop[0] != 0) g.bins.erase({pl.nf[0], pl.nf[1], pl.nf[2], pl.w, 1, drop[0], drop[1], drop[2]});
        }
    } catch (...) {
        cx.timer.on = timing;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        cx.bin_tune.erase(key);
        throw;
    }
    cx.timer.on = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    LAM_REQUIRE(best[0] != 0, LAM_ERR_ARG, "no spreading tile shape fits this fine grid");
    cx.bin_tune[key] = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = best;
    }
    if (best_ms) *best_ms = best_t;
    return best;
}

}  // namespace lam
