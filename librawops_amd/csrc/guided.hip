// Working-Y guided filter for gfx950 (SpatialOps.cpp:159-317), box-moment
// formulation with true-edge actual counts.
//
// Algebra.  The reference's ordered centred form
//   q_i = p_i + mean_{k in W_i} [ (p_k - p_i) + mp_k + a_k (Y_i - Y_k - mg_k) ]
// with mp_k = mean_{j in W_k}(p_j - p_k), mg_k = mean(Y_j - Y_k) and
// a_k = cov_k / (var_k + eps) equals  q_i = mean_{k in W_i}(a_k) Y_i +
// mean_{k in W_i}(b_k),  b_k = mean_{W_k}(p) - a_k mean_{W_k}(Y).
// Both box means are computed with sliding vertical sums and per-row
// wavefront prefix scans (O(1) work per pixel for any radius), in FP64.  The
// moments are accumulated on values shifted by a per-strip constant so that
// E[Y^2] - E[Y]^2 does not cancel catastrophically.
//
// Two strip kernels: `coefficients` writes a_k, b_k for every centre of the
// viewport grown by r (computed in FP64, stored as FP32 planes: the stored
// values carry a relative error of 2^-24, far inside the filter's 1e-5
// tolerance, and halve the workspace traffic); `apply` box-averages them in
// FP64, forms q and runs the fused post-filter op program (exposure, curves).
#include "device_common.hpp"
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kStripThreads = 256;
constexpr int kCols = 1;                          // columns per thread
constexpr int kStripCols = kStripThreads * kCols; // columns per strip (centres + 2r halo)
constexpr int kBandRows = 32;                     // band height when the grid fills the GPU
constexpr int kMinBandRows = 8;

// Rows per band: kBandRows, halved (down to kMinBandRows) while the launch
// would have fewer than ~4 workgroups per CU; a shorter band repeats the
// 2r-row warm-up of the vertical window but keeps small images parallel.
inline int band_rows(int strips, int rows) {
    int band = kBandRows;
    while (band > kMinBandRows && (long)strips * ((rows + band - 1) / band) < 1024) band /= 2;
    return band;
}

struct GuidedArgs {
    int band;                // rows per band (blockIdx.y)
    uint32_t fused;          // post ops [0, fused) run in gf_apply, [fused, n) in gf_ops_tail
    const float* src;
    uint32_t src_stride;
    int W, H;                // native raster size
    int mip;
    int LW, LH;              // level image size
    int cx0, cy0, cw, ch;    // centre region (level coords) = viewport grown by r, clipped
    int vx, vy, vw, vh;      // viewport (level coords)
    int r;
    double eps, wr, wb;
    float* coef;             // 6 planes [ch][cw] (a0 a1 a2 b0 b1 b2), plane stride cw * ch
    float* out;
    uint32_t ostride;
    uint32_t* flags;
    const OpList* __restrict__ post;
    size_t zsrc, zcoef, zout;   // floats between consecutive images of a stack (blockIdx.z)
};

// Image blockIdx.z of a stack of equal-size images, consecutive in memory.
__device__ __forceinline__ GuidedArgs slice_of(const GuidedArgs& g0) {
    GuidedArgs g = g0;
    const size_t z = blockIdx.z;
    g.src += z * g0.zsrc;
    g.out += z * g0.zout;
    if (g0.coef) g.coef += z * g0.zcoef;
    return g;
}

// Raster pixel at level coordinates; mip 1/2 average clipped native
// footprints in FP64 (RasterSourceNode::render_level RawEngine.cpp:363-406).
__device__ __forceinline__ void level_pixel(const GuidedArgs& g, int x, int y, float p[3]) {
    if (g.mip == 0) {
        const float* q = g.src + ((size_t)y * g.src_stride + (size_t)x) * 3;
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        return;
    }
    const int s = 1 << g.mip;
    const int x0 = x * s, y0 = y * s;
    const int x1 = min(x0 + s, g.W), y1 = min(y0 + s, g.H);
    double sum[3] = {0, 0, 0};
    for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) {
            const float* q = g.src + ((size_t)yy * g.src_stride + (size_t)xx) * 3;
            sum[0] += q[0]; sum[1] += q[1]; sum[2] += q[2];
        }
    const double count = (double)((x1 - x0) * (y1 - y0));
    p[0] = (float)(sum[0] / count); p[1] = (float)(sum[1] / count); p[2] = (float)(sum[2] / count);
}

// Working Y of an RGB triple: (g + wr(r - g)) + wb(b - g) (SpatialOps.cpp:213-219).
__device__ __forceinline__ double working_y(const GuidedArgs& g, double r, double gg, double b) {
    return (gg + g.wr * (r - gg)) + g.wb * (b - gg);
}

// Box sums over a strip of kStripCols columns and a band of rows, O(1) work
// per pixel for any radius.  Thread t owns columns strip_x - r + t kCols + j
// (j < kCols) and keeps their vertical window sums (rows k-r..k+r, clipped)
// of the N values src(col, row) produces, sliding them down one row at a
// time.  Per row the strip's inclusive prefix along x is formed (serial over
// a thread's columns, then a wavefront shuffle scan of the thread totals and
// the wave totals through LDS), so each centre's (2r+1)^2 box is two prefix
// reads; sink(x, row, S, count) receives the clipped box sums and the true
// window size.  Centres x in [strip_x, min(xend, strip_x + kStripCols - 2r))
// are produced, kCols per thread.  Windows are clipped to [0, LW) x [0, LH);
// src is only asked for columns in [c_lo, c_hi), which must hold every
// clipped window of those centres.
template <int N, class Src, class Sink>
__device__ __forceinline__ void strip_box(int LW, int LH, int c_lo, int c_hi, int r, int strip_x, int xend, int y0, int y1,
                                          double* prefix /* [N][kStripCols] */, double* totals /* [4][N] */,
                                          Src src, Sink sink) {
    constexpr int K = kCols;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int span = kStripCols - 2 * r;
    const int base = strip_x - r;                 // column of prefix index 0
    double vs[K][N];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int m = 0; m < N; ++m) vs[j][m] = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int col = base + t * K + j;
        if (col >= c_lo && col < c_hi)
            for (int y = max(0, y0 - r); y < min(LH, y0 + r + 1); ++y) {
                double v[N];
                src(col, y, v);
#pragma unroll
                for (int m = 0; m < N; ++m) vs[j][m] += v[m];
            }
    }
    for (int k = y0; k < y1; ++k) {
        if (k > y0) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int col = base + t * K + j;
                if (col < c_lo || col >= c_hi) continue;
                if (k + r < LH) {
                    double v[N];
                    src(col, k + r, v);
#pragma unroll
                    for (int m = 0; m < N; ++m) vs[j][m] += v[m];
                }
                if (k - r - 1 >= 0) {
                    double v[N];
                    src(col, k - r - 1, v);
#pragma unroll
                    for (int m = 0; m < N; ++m) vs[j][m] -= v[m];
                }
            }
        }
        // Inclusive prefix along the strip.
        double run[N], inc[N];
#pragma unroll
        for (int m = 0; m < N; ++m) run[m] = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
            for (int m = 0; m < N; ++m) run[m] += vs[j][m];
#pragma unroll
        for (int m = 0; m < N; ++m) inc[m] = run[m];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1)
#pragma unroll
            for (int m = 0; m < N; ++m) {
                const double o = __shfl_up(inc[m], d, 64);
                if (lane >= d) inc[m] += o;
            }
        if (lane == 63)
#pragma unroll
            for (int m = 0; m < N; ++m) totals[wave * N + m] = inc[m];
        __syncthreads();
        double acc[N];
#pragma unroll
        for (int m = 0; m < N; ++m) acc[m] = inc[m] - run[m];   // exclusive within the wave
        for (int w = 0; w < wave; ++w)
#pragma unroll
            for (int m = 0; m < N; ++m) acc[m] += totals[w * N + m];
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
            for (int m = 0; m < N; ++m) {
                acc[m] += vs[j][m];
                prefix[m * kStripCols + t * K + j] = acc[m];
            }
        __syncthreads();
        const int rows = min(LH, k + r + 1) - max(0, k - r);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int o = t * K + j;              // centre index within the strip
            const int x = strip_x + o;
            if (o >= span || x >= xend) continue;
            const int lo = max(0, x - r) - base;
            const int hi = min(LW, x + r + 1) - base;
            double S[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
                S[m] = prefix[m * kStripCols + hi - 1] - (lo > 0 ? prefix[m * kStripCols + lo - 1] : 0.0);
            sink(x, k, S, (double)((hi - lo) * rows));
        }
        __syncthreads();
    }
}

// Moments of one (shifted) pixel: Y, p0, p1, p2, YY, Yp0, Yp1, Yp2.
__device__ __forceinline__ void moments(const GuidedArgs& g, const float p[3], const double shift[3],
                                        double m[8], uint32_t& err) {
    if (!finite(p[0]) || !finite(p[1]) || !finite(p[2])) err |= FLAG_INVALID;
    const double d0 = (double)p[0] - shift[0], d1 = (double)p[1] - shift[1], d2 = (double)p[2] - shift[2];
    const double y = working_y(g, d0, d1, d2);
    m[0] = y; m[1] = d0; m[2] = d1; m[3] = d2;
    m[4] = y * y; m[5] = y * d0; m[6] = y * d1; m[7] = y * d2;
}

__global__ void __launch_bounds__(kStripThreads) gf_coefficients(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    __shared__ double prefix[kStripCols * 8];
    __shared__ double totals[4 * 8];
    const int span = kStripCols - 2 * g.r;
    const int strip_x = g.cx0 + (int)blockIdx.x * span;
    const int k0 = g.cy0 + (int)blockIdx.y * g.band;
    const int k1 = min(k0 + g.band, g.cy0 + g.ch);
    uint32_t err = 0;
    // Per-strip shift: the first centre pixel of the strip.
    double shift[3];
    {
        float p[3];
        level_pixel(g, min(strip_x, g.LW - 1), k0, p);
        shift[0] = p[0]; shift[1] = p[1]; shift[2] = p[2];
    }
    const double shift_y = working_y(g, shift[0], shift[1], shift[2]);
    auto src = [&](int x, int y, double m[8]) {
        float p[3];
        level_pixel(g, x, y, p);
        moments(g, p, shift, m, err);
    };
    auto sink = [&](int kx, int k, const double s[8], double n) {
        const double inv_n = 1.0 / n;
        const double my = s[0] * inv_n;
        double var = s[4] * inv_n - my * my;
        if (var < 0) var = 0;
        const double inv_d = 1.0 / (var + g.eps);
        const size_t plane = (size_t)g.cw * g.ch;
        float* c = g.coef + (size_t)(k - g.cy0) * g.cw + (size_t)(kx - g.cx0);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const double mp = s[1 + ch] * inv_n;
            const double cov = s[5 + ch] * inv_n - my * mp;
            const double a = cov * inv_d;
            // b in unshifted coordinates: mean(p) - a mean(Y).
            const double b = (mp - a * my) + (shift[ch] - a * shift_y);
            if (!__builtin_isfinite(a) || !__builtin_isfinite(b)) err |= FLAG_OVERFLOW;
            c[ch * plane] = (float)a;
            c[(3 + ch) * plane] = (float)b;
        }
    };
    strip_box<8>(g.LW, g.LH, 0, g.LW, g.r, strip_x, g.cx0 + g.cw, k0, k1, prefix, totals, src, sink);
    if (err && g.flags) atomicOr(g.flags, err);
}

__global__ void __launch_bounds__(kStripThreads) gf_apply(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    __shared__ double prefix[kStripCols * 6];
    __shared__ double totals[4 * 6];
    const int span = kStripCols - 2 * g.r;
    const int strip_x = g.vx + (int)blockIdx.x * span;
    const int i0 = g.vy + (int)blockIdx.y * g.band;
    const int i1 = min(i0 + g.band, g.vy + g.vh);
    uint32_t err = 0;
    // Coefficients exist for centres in [cx0, cx0 + cw) x [cy0, cy0 + ch):
    // exactly the clipped windows of the viewport's pixels.
    const size_t plane = (size_t)g.cw * g.ch;
    auto src = [&](int k_x, int k_y, double c[6]) {
        const float* q = g.coef + (size_t)(k_y - g.cy0) * g.cw + (size_t)(k_x - g.cx0);
#pragma unroll
        for (int m = 0; m < 6; ++m) c[m] = (double)q[m * plane];
    };
    auto sink = [&](int x, int i, const double S[6], double n) {
        const double inv_n = 1.0 / n;
        float p[3];
        level_pixel(g, x, i, p);
        const double y = working_y(g, (double)p[0], (double)p[1], (double)p[2]);
        float o[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const double q = (S[ch] * inv_n) * y + S[3 + ch] * inv_n;
            const double delta = q - (double)p[ch];
            if (!__builtin_isfinite(delta)) err |= FLAG_OVERFLOW;
            if (delta == 0) { o[ch] = p[ch]; continue; }   // exact copy, as the reference
            const double value = (double)p[ch] + delta;
            if (value < -3.4028234663852886e38 || value > 3.4028234663852886e38) err |= FLAG_OVERFLOW;
            o[ch] = (float)value;
        }
        apply_ops(*g.post, 0, g.fused, o[0], o[1], o[2], err);
        float* d = g.out + ((size_t)(i - g.vy) * g.ostride + (size_t)(x - g.vx)) * 3;
        d[0] = o[0]; d[1] = o[1]; d[2] = o[2];
    };
    strip_box<6>(g.LW, g.LH, g.cx0, g.cx0 + g.cw, g.r, strip_x, g.vx + g.vw, i0, i1, prefix, totals, src, sink);
    if (err && g.flags) atomicOr(g.flags, err);
}

// The post-filter ops from g.fused on (g.fused is a curve op), in place over
// the filtered output.  The curve op's knot tables are staged in LDS when
// every curve has at most kTailKnots knots: its lookups are data-dependent,
// and from LDS they do not wait on divergent global loads.
constexpr int kTailKnots = 64;
__global__ void __launch_bounds__(256) gf_ops_tail(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    __shared__ double tables[4 * 5 * kTailKnots];
    const size_t n = (size_t)g.vw * g.vh;
    Const<OpList>& L = as_const(*g.post);
    const uint32_t end = L.n;
    const DevCurves* cv = L.op[g.fused].curves;
    bool staged = true;
    for (int c = 0; c < 4; ++c) staged = staged && (as_const(*cv).c[c].identity || as_const(*cv).c[c].n <= kTailKnots);
    LdsCurve curves[4];
    const uint32_t cubic = as_const(*cv).cubic;
    if (staged) {
        for (int c = 0; c < 4; ++c) curves[c] = stage_curve(&cv->c[c], tables + c * 5 * kTailKnots);
        __syncthreads();
    }
    uint32_t err = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / g.vw), x = (int)(i - (size_t)y * g.vw);
        float* d = g.out + ((size_t)y * g.ostride + (size_t)x) * 3;
        float p[3] = {d[0], d[1], d[2]};
        if (staged) {
            for (int c = 0; c < 3; ++c) p[c] = curves_channel_lds(curves, cubic, c, p[c], err);
            apply_ops(*g.post, g.fused + 1, end, p[0], p[1], p[2], err);
        } else {
            apply_ops(*g.post, g.fused, end, p[0], p[1], p[2], err);
        }
        d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
    }
    if (err && g.flags) atomicOr(g.flags, err);
}

// Radius 0: the filter is the identity (SpatialOps.cpp:180); only the
// post-filter program runs.
__global__ void gf_identity(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    const size_t n = (size_t)g.vw * g.vh;
    uint32_t err = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / g.vw), x = (int)(i - (size_t)y * g.vw);
        float p[3];
        level_pixel(g, g.vx + x, g.vy + y, p);
        if (!finite(p[0]) || !finite(p[1]) || !finite(p[2])) err |= FLAG_INVALID;
        apply_ops(*g.post, p[0], p[1], p[2], err);
        float* d = g.out + ((size_t)y * g.ostride + (size_t)x) * 3;
        d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
    }
    if (err && g.flags) atomicOr(g.flags, err);
}


// ------------------------------------------------------------- backward --
//
// Reverse mode of q_i = mean_{k in W_i}(a_k) Y_i + mean_{k in W_i}(b_k) with
// the per-centre statistics recomputed from the input (not stored by the
// forward).  Pass 1 (per centre k): box sums of the 8 input moments and of
// gq_i Y_i / n_i, gq_i / n_i give a_k, b_k and their adjoints, pushed back to
// the moments: P_k = dmu_p / n_k, Q_k = dE[Yp] / n_k, R_k = dmu_Y / n_k,
// S_k = dE[Y^2] / n_k (and a_k for the direct Y term).  Pass 2 (per pixel
// j): box sums of those 11 values give dL/dp_j.  Moments are unshifted FP64;
// the per-centre terms are stored as FP32 planes.

struct GuidedBackArgs {
    int band;                // rows per band (blockIdx.y)
    const float* p;          // raster, level 0 (W x H, stride)
    uint32_t stride;
    int W, H, r;
    double eps, wr, wb;
    const float* gq;         // dL/dq over the viewport
    uint32_t gstride;
    int vx, vy, vw, vh;
    float* ws;               // 11 planes [H][W] of per-centre adjoint terms
    float* gp;               // dL/dp, W x H x 3 (written in full)
    double* geps;            // accumulated dL/deps
    size_t zp, zgq, zws, zgp;   // floats between consecutive images of a stack (blockIdx.z)
};

__device__ __forceinline__ GuidedBackArgs slice_of(const GuidedBackArgs& g0) {
    GuidedBackArgs g = g0;
    const size_t z = blockIdx.z;
    g.p += z * g0.zp;
    g.gq += z * g0.zgq;
    g.ws += z * g0.zws;
    g.gp += z * g0.zgp;
    return g;
}

__device__ __forceinline__ double window_count(int x, int y, int W, int H, int r) {
    return (double)((min(W, x + r + 1) - max(0, x - r)) * (min(H, y + r + 1) - max(0, y - r)));
}

__global__ void __launch_bounds__(kStripThreads) gf_back_centres(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double prefix[kStripCols * 14];
    __shared__ double totals[4 * 14];
    __shared__ double eps_part[kStripThreads / 64];
    const int span = kStripCols - 2 * g.r;
    const int strip_x = (int)blockIdx.x * span;
    const int y0 = (int)blockIdx.y * g.band, y1 = min(y0 + g.band, g.H);
    double geps = 0.0;
    auto src = [&](int x, int y, double v[14]) {
        const float* q = g.p + ((size_t)y * g.stride + (size_t)x) * 3;
        const double pr = q[0], pg = q[1], pb = q[2];
        const double Y = (pg + g.wr * (pr - pg)) + g.wb * (pb - pg);
        v[0] = Y; v[1] = pr; v[2] = pg; v[3] = pb;
        v[4] = Y * Y; v[5] = Y * pr; v[6] = Y * pg; v[7] = Y * pb;
        const int lx = x - g.vx, ly = y - g.vy;
        if (lx >= 0 && ly >= 0 && lx < g.vw && ly < g.vh) {
            const float* d = g.gq + ((size_t)ly * g.gstride + (size_t)lx) * 3;
            const double inv_n = 1.0 / window_count(x, y, g.W, g.H, g.r);
#pragma unroll
            for (int c = 0; c < 3; ++c) { v[8 + c] = (double)d[c] * Y * inv_n; v[11 + c] = (double)d[c] * inv_n; }
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) v[8 + c] = 0.0;
        }
    };
    auto sink = [&](int x, int y, const double S[14], double n) {
        const double inv_n = 1.0 / n;
        const double my = S[0] * inv_n, eyy = S[4] * inv_n;
        const double var = eyy - my * my, D = (var < 0 ? 0 : var) + g.eps;
        const double inv_D = 1.0 / D;
        double a[3], mp[3], ga[3], gmp[3], gcov[3];
        double gmy = 0, gD = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            mp[c] = S[1 + c] * inv_n;
            const double cov = S[5 + c] * inv_n - my * mp[c];
            a[c] = cov * inv_D;
            const double GA = S[8 + c], GB = S[11 + c];
            gmp[c] = GB;
            ga[c] = GA - my * GB;
            gmy -= a[c] * GB;
            gcov[c] = ga[c] * inv_D;
            gD -= ga[c] * a[c] * inv_D;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gmy -= gcov[c] * mp[c];
            gmp[c] -= gcov[c] * my;
        }
        gmy -= 2.0 * my * gD;
        geps += gD;
        const size_t plane = (size_t)g.W * g.H;
        float* w = g.ws + (size_t)y * g.W + (size_t)x;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            w[c * plane] = (float)(gmp[c] * inv_n);
            w[(3 + c) * plane] = (float)(gcov[c] * inv_n);
            w[(8 + c) * plane] = (float)a[c];
        }
        w[6 * plane] = (float)(gmy * inv_n);
        w[7 * plane] = (float)(gD * inv_n);
    };
    strip_box<14>(g.W, g.H, 0, g.W, g.r, strip_x, g.W, y0, y1, prefix, totals, src, sink);
    for (int off = 32; off > 0; off >>= 1) geps += __shfl_xor(geps, off, 64);
    if ((threadIdx.x & 63) == 0) eps_part[threadIdx.x >> 6] = geps;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kStripThreads / 64; ++w) t += eps_part[w];
        atomicAdd(g.geps, t);
    }
}

__global__ void __launch_bounds__(kStripThreads) gf_back_pixels(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double prefix[kStripCols * 11];
    __shared__ double totals[4 * 11];
    const int span = kStripCols - 2 * g.r;
    const int strip_x = (int)blockIdx.x * span;
    const int y0 = (int)blockIdx.y * g.band, y1 = min(y0 + g.band, g.H);
    const size_t plane = (size_t)g.W * g.H;
    auto src = [&](int x, int y, double v[11]) {
        const float* w = g.ws + (size_t)y * g.W + (size_t)x;
#pragma unroll
        for (int m = 0; m < 11; ++m) v[m] = (double)w[m * plane];
    };
    auto sink = [&](int x, int y, const double S[11], double n) {
        const float* q = g.p + ((size_t)y * g.stride + (size_t)x) * 3;
        const double p[3] = {q[0], q[1], q[2]};
        const double Y = (p[1] + g.wr * (p[0] - p[1])) + g.wb * (p[2] - p[1]);
        double gY = S[6] + 2.0 * S[7] * Y;
        double gp[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gp[c] = S[c] + S[3 + c] * Y;
            gY += S[3 + c] * p[c];
        }
        const int lx = x - g.vx, ly = y - g.vy;
        if (lx >= 0 && ly >= 0 && lx < g.vw && ly < g.vh) {
            const float* d = g.gq + ((size_t)ly * g.gstride + (size_t)lx) * 3;
            const double inv_n = 1.0 / n;
#pragma unroll
            for (int c = 0; c < 3; ++c) gY += (double)d[c] * (S[8 + c] * inv_n);
        }
        gp[0] += gY * g.wr;
        gp[1] += gY * (1.0 - g.wr - g.wb);
        gp[2] += gY * g.wb;
        float* o = g.gp + ((size_t)y * g.W + (size_t)x) * 3;
        o[0] = (float)gp[0]; o[1] = (float)gp[1]; o[2] = (float)gp[2];
    };
    strip_box<11>(g.W, g.H, 0, g.W, g.r, strip_x, g.W, y0, y1, prefix, totals, src, sink);
}

}  // namespace

size_t guided_workspace_bytes(const GuidedLaunch& L) {
    const int s = 1 << L.mip;
    const int LW = ((int)L.width + s - 1) / s, LH = ((int)L.height + s - 1) / s;
    const int r = (int)L.radius;
    const int cx0 = std::max(0, (int)L.viewport.x - r), cy0 = std::max(0, (int)L.viewport.y - r);
    const int cx1 = std::min(LW, (int)(L.viewport.x + L.viewport.width) + r);
    const int cy1 = std::min(LH, (int)(L.viewport.y + L.viewport.height) + r);
    return (size_t)(cx1 - cx0) * (size_t)(cy1 - cy0) * 6 * sizeof(float) * std::max(1u, L.slices);
}

size_t guided_backward_workspace_bytes(uint32_t width, uint32_t height) {
    return (size_t)width * height * 11 * sizeof(float);
}

hipError_t launch_guided_backward(const GuidedBackwardLaunch& L, hipStream_t stream) {
    GuidedBackArgs g{};
    g.p = L.src;
    g.stride = L.src_stride;
    g.W = (int)L.width;
    g.H = (int)L.height;
    g.r = (int)L.radius;
    g.eps = L.epsilon;
    g.wr = L.wr;
    g.wb = L.wb;
    g.gq = L.g_out;
    g.gstride = L.g_stride;
    g.vx = (int)L.viewport.x; g.vy = (int)L.viewport.y;
    g.vw = (int)L.viewport.width; g.vh = (int)L.viewport.height;
    g.ws = static_cast<float*>(L.workspace);
    g.gp = L.g_in;
    g.geps = L.g_eps;
    const unsigned nz = std::max(1u, L.slices);
    g.zp = L.src_zstride;
    g.zgq = L.g_out_zstride;
    g.zws = (size_t)11 * g.W * g.H;
    g.zgp = (size_t)3 * g.W * g.H;
    const int span = kStripCols - 2 * g.r;
    const int strips = (g.W + span - 1) / span;
    g.band = band_rows(strips, g.H);   // as for one image: the same sums, bit for bit
    dim3 grid((unsigned)strips, (unsigned)((g.H + g.band - 1) / g.band), nz);
    hipLaunchKernelGGL(gf_back_centres, grid, dim3(kStripThreads), 0, stream, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gf_back_pixels, grid, dim3(kStripThreads), 0, stream, g);
    return hipGetLastError();
}

hipError_t launch_guided(const GuidedLaunch& L, const OpList* post, hipStream_t stream) {
    GuidedArgs g{};
    g.src = L.src;
    g.src_stride = L.src_stride;
    g.W = (int)L.width;
    g.H = (int)L.height;
    g.mip = (int)L.mip;
    const int s = 1 << g.mip;
    g.LW = (g.W + s - 1) / s;
    g.LH = (g.H + s - 1) / s;
    g.r = (int)L.radius;
    g.eps = L.epsilon;
    g.wr = L.wr;
    g.wb = L.wb;
    g.vx = (int)L.viewport.x; g.vy = (int)L.viewport.y;
    g.vw = (int)L.viewport.width; g.vh = (int)L.viewport.height;
    g.cx0 = std::max(0, g.vx - g.r); g.cy0 = std::max(0, g.vy - g.r);
    g.cw = std::min(g.LW, g.vx + g.vw + g.r) - g.cx0;
    g.ch = std::min(g.LH, g.vy + g.vh + g.r) - g.cy0;
    g.out = L.out;
    g.ostride = L.ostride;
    g.flags = L.flags;
    g.post = post;
    const unsigned nz = std::max(1u, L.slices);
    g.zsrc = L.src_zstride;
    g.zout = L.out_zstride;
    if (!g.vw || !g.vh) return hipSuccess;
    if (g.r == 0) {
        const size_t n = (size_t)g.vw * g.vh;
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535);
        hipLaunchKernelGGL(gf_identity, dim3(blocks, 1, nz), dim3(256), 0, stream, g);
        return hipGetLastError();
    }
    g.coef = static_cast<float*>(L.workspace);
    g.zcoef = (size_t)6 * g.cw * g.ch;
    const int span = kStripCols - 2 * g.r;
    const int strips_c = (g.cw + span - 1) / span;
    g.band = band_rows(strips_c, g.ch);   // as for one image: the same sums, bit for bit
    dim3 grid_c((unsigned)strips_c, (unsigned)((g.ch + g.band - 1) / g.band), nz);
    hipLaunchKernelGGL(gf_coefficients, grid_c, dim3(kStripThreads), 0, stream, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int strips_a = (g.vw + span - 1) / span;
    g.band = band_rows(strips_a, g.vh);
    dim3 grid_a((unsigned)strips_a, (unsigned)((g.vh + g.band - 1) / g.band), nz);
    g.fused = L.fused_ops;
    hipLaunchKernelGGL(gf_apply, grid_a, dim3(kStripThreads), 0, stream, g);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (L.fused_ops < L.post_ops) {
        const size_t n = (size_t)g.vw * g.vh;
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535);
        hipLaunchKernelGGL(gf_ops_tail, dim3(blocks, 1, nz), dim3(256), 0, stream, g);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace rawops
