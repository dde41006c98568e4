// Working-Y guided filter for gfx950 (SpatialOps.cpp:159-317), box-moment
// formulation with true-edge actual counts.
//
// Algebra.  The reference's ordered centred form
//   q_i = p_i + mean_{k in W_i} [ (p_k - p_i) + mp_k + a_k (Y_i - Y_k - mg_k) ]
// with mp_k = mean_{j in W_k}(p_j - p_k), mg_k = mean(Y_j - Y_k) and
// a_k = cov_k / (var_k + eps) equals  q_i = mean_{k in W_i}(a_k) Y_i +
// mean_{k in W_i}(b_k),  b_k = mean_{W_k}(p) - a_k mean_{W_k}(Y).
// Both box means are sums over strips of pixel columns: vertical window sums
// slid down a band of rows per column, then sums along x over each centre's
// window (O(1) work per pixel for any radius), in FP64.  The moments are
// accumulated on values shifted by a per-strip constant so that
// E[Y^2] - E[Y]^2 does not cancel catastrophically.
//
// Two strip kernels: `coefficients` writes a_k, b_k for every centre of the
// viewport grown by r (computed in FP64, stored as FP32 planes: the stored
// values carry a relative error of 2^-24, far inside the filter's 1e-5
// tolerance); `apply` box-averages them in FP64, forms q and runs the
// post-filter op program (exposure, curves).  For r <= 16 the sums along x
// are sliding sums over runs of 16 centres in LDS (strip_runs); larger radii
// use per-row wavefront prefix scans (strip_box).
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
    uint32_t fused;          // post ops [0, fused) run in the apply pass, the rest (from a curve op) after
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
    __shared__ __attribute__((aligned(16))) double tables[4 * poly_table_doubles(kTailKnots)];
    const size_t n = (size_t)g.vw * g.vh;
    Const<OpList>& L = as_const(*g.post);
    const uint32_t end = L.n;
    const DevCurves* cv = L.op[g.fused].curves;
    bool staged = true;
    for (int c = 0; c < 4; ++c) staged = staged && (as_const(*cv).c[c].identity || as_const(*cv).c[c].n <= kTailKnots);
    LdsPoly curves[4];
    const uint32_t cubic = as_const(*cv).cubic;
    if (staged) {
        for (int c = 0; c < 4; ++c) curves[c] = stage_poly(&cv->c[c], cubic, tables + c * poly_table_doubles(kTailKnots));
        __syncthreads();
    }
    uint32_t err = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / g.vw), x = (int)(i - (size_t)y * g.vw);
        float* d = g.out + ((size_t)y * g.ostride + (size_t)x) * 3;
        float p[3] = {d[0], d[1], d[2]};
        if (staged) {
            curves3_lds(curves, p, err);
            apply_ops(*g.post, g.fused + 1, end, p[0], p[1], p[2], err);
        } else {
            apply_ops(*g.post, g.fused, end, p[0], p[1], p[2], err);
        }
        d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
    }
    if (err && g.flags) atomicOr(g.flags, err);
}

// Box sums for a compile-time radius R (1..16) without per-row prefix scans.
// The strip is kRunCols = 256 pixel columns x = strip_x - R + t; thread t
// keeps the vertical window sums (rows k-R..k+R, clipped) of the N values of
// its column, slid down one row per step, with the row data of a step loaded
// two steps ahead (fetch(x, y) returns it, eval(data, x, y, v) turns it into
// the N values; fetch must be branch-free, so that the compiler's count of
// loads in flight stays exact and the loads overlap the steps).  Per row:
//   1. thread t writes its N sums into slot t of N LDS rows zero-padded by R;
//   2. thread t = 16 m + j (m < N) forms the sums along x of value m for the
//      16 centres 16 j .. 16 j + 15 as sliding sums (2R + 16 reads) and, once
//      every lane of its wavefront has read (a wavefront fence orders the
//      reads before the writes), writes the sum of centre c back into slot c
//      of row m; a wavefront touches only the rows of its own values;
//   3. thread t, the centre x = strip_x - R + t for t in [R, 256 - R), reads
//      its N box sums from slot t and calls sink(data, x, k, S, cols, rows),
//      data being what centre(x, k) loaded a step earlier, cols x rows the
//      clipped window.
// Slot t is written in step 1 and read in step 3 only by thread t, so two
// barriers per row suffice.  A row holds logical index i = pixel + R at
// run_slot(i) (one pad double per 16, so the 16 lanes of a value read
// distinct banks); rows are plane doubles apart, plane = 16 mod 32, so the 4
// values of a wavefront spread over the banks as well.  Windows are clipped
// to [0, LW) x [0, LH); only pixels of the data region [c_lo, c_hi) x
// [r_lo, r_hi) enter the sums (it must hold every clipped window of the
// band's centres), and fetch and centre are called only inside it
// (coordinates clamped into it), so their loads are unconditional.
constexpr int kRunCols = 256;
constexpr int kRunThreads = 256;
constexpr int kRunLen = 16;   // centres per lane in the sums along x

__device__ __forceinline__ int run_slot(int i) { return i + (i >> 4); }

template <int R> struct RunGeom {
    static constexpr int logical = kRunCols + 2 * R;                      // zero-padded row
    static constexpr int used = logical + (logical >> 4) + 1;
    static constexpr int plane = ((used + 15) / 32) * 32 + 16;            // >= used, = 16 mod 32
    static_assert(plane >= used, "run plane");
};

// s[t] = sum_{i=0}^{2R} b[run_slot(i0 + t + i)], t < kRunLen, as sliding sums:
// the first window as two partial sums (so the adds overlap), then the
// entering-minus-leaving differences, formed independently of each other, so
// the dependent chain is one add per centre.
template <int R>
__device__ __forceinline__ void run_sums16(const double* __restrict__ b, int i0, double s[kRunLen]) {
    double e = 0.0, o = 0.0;
#pragma unroll
    for (int i = 0; i <= 2 * R; i += 2) e += b[run_slot(i0 + i)];
#pragma unroll
    for (int i = 1; i <= 2 * R; i += 2) o += b[run_slot(i0 + i)];
    double d[kRunLen];
#pragma unroll
    for (int t = 1; t < kRunLen; ++t) d[t] = b[run_slot(i0 + 2 * R + t)] - b[run_slot(i0 + t - 1)];
    double acc = e + o;
    s[0] = acc;
#pragma unroll
    for (int t = 1; t < kRunLen; ++t) {
        acc += d[t];
        s[t] = acc;
    }
}

template <int N, int R, int RP, class Fetch, class Eval, class Centre, class Sink>
__device__ __forceinline__ void strip_runs(int LW, int LH, int c_lo, int c_hi, int r_lo, int r_hi, int strip_x,
                                           int xend, int y0, int y1, double* rows /* [RP][N][plane] */, Fetch fetch,
                                           Eval eval, Centre centre, Sink sink) {
    static_assert(RP == 1 || RP == 2, "rows per step");
    constexpr int P = RunGeom<R>::plane;
    constexpr int kRuns = kRunCols / kRunLen;          // runs of 16 centres per value and row
    constexpr int kTasks = RP * N * kRuns;             // (row, value, run) sum tasks per step
    constexpr int kPasses = (kTasks + kRunThreads - 1) / kRunThreads;
    const int t = threadIdx.x;
    const int x = strip_x - R + t;
    const bool own = x >= c_lo && x < c_hi;
    const int xc = min(max(x, c_lo), c_hi - 1);
    const int slot = run_slot(t + R);
    auto row_of = [&](int y) { return min(max(y, r_lo), r_hi - 1); };
    // Rows that enter the sums: the image rows of the data region (every
    // clipped window of the band's centres lies inside it).
    const int row_lo = max(0, r_lo), row_hi = min(LH, r_hi);
    for (int i = t; i < RP * N * P; i += kRunThreads) rows[i] = 0.0;
    // Warm-up: the window of row y0 - 1, four rows of loads in flight.
    double V[N];
#pragma unroll
    for (int m = 0; m < N; ++m) V[m] = 0.0;
    for (int y = y0 - 1 - R; y <= y0 - 1 + R; y += 4) {
        decltype(fetch(0, 0)) d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) d[u] = fetch(xc, row_of(y + u));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int yy = y + u;
            if (own && yy <= y0 - 1 + R && yy >= row_lo && yy < row_hi) {
                double v[N];
                eval(d[u], x, yy, v);
#pragma unroll
                for (int m = 0; m < N; ++m) V[m] += v[m];
            }
        }
    }
    const bool is_centre = t >= R && t < kRunCols - R && x < xend;
    // Window of row k: add row k + R, drop row k - R - 1 (data in en / lv),
    // then publish the N sums in slot t of the LDS rows at dst.
    auto slide = [&](int k, const auto& en, const auto& lv, double* dst) {
        if (own) {
            double v[N];
            if (k + R < row_hi) {
                eval(en, x, k + R, v);
#pragma unroll
                for (int m = 0; m < N; ++m) V[m] += v[m];
            }
            if (k - R - 1 >= row_lo) {
                eval(lv, x, k - R - 1, v);
#pragma unroll
                for (int m = 0; m < N; ++m) V[m] -= v[m];
            }
        }
#pragma unroll
        for (int m = 0; m < N; ++m) dst[m * P + slot] = V[m];
    };
    // Sums along x of every published row: task u = 16 (RP-row value index)
    // + run; the 16 tasks of one LDS row are 16 consecutive lanes of one
    // wavefront, so a wavefront fence orders all reads of a row before its
    // in-place writes.
    auto runs = [&]() {
#pragma unroll
        for (int pass = 0; pass < kPasses; ++pass) {
            const int u = t + pass * kRunThreads;
            if (kTasks % kRunThreads == 0 || u < kTasks) {
                double* b = rows + (u / kRuns) * P;
                const int vj = u % kRuns;
                double Hs[kRunLen];
                run_sums16<R>(b, kRunLen * vj, Hs);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int i = 0; i < kRunLen; ++i) b[run_slot(kRunLen * vj + i + R)] = Hs[i];
            }
        }
    };
    auto finish = [&](int k, const double* src, const auto& cd) {
        if (is_centre) {
            double S[N];
#pragma unroll
            for (int m = 0; m < N; ++m) S[m] = src[m * P + slot];
            const int cols = min(LW, x + R + 1) - max(0, x - R);
            const int rws = min(LH, k + R + 1) - max(0, k - R);
            sink(cd, x, k, S, cols, rws);
        }
    };
    if constexpr (RP == 1) {
        // One row per step; steps alternate between two sets of buffers (no
        // register copies, so the loads of the next two steps stay in flight).
        auto step = [&](int k, auto& en, auto& lv, auto& cd) {
            slide(k, en, lv, rows);
            en = fetch(xc, row_of(k + 2 + R));
            lv = fetch(xc, row_of(k + 1 - R));
            __syncthreads();
            runs();
            __syncthreads();
            finish(k, rows, cd);
            cd = centre(xc, row_of(k + 2));
        };
        auto en0 = fetch(xc, row_of(y0 + R));
        auto lv0 = fetch(xc, row_of(y0 - R - 1));
        auto cd0 = centre(xc, row_of(y0));
        auto en1 = fetch(xc, row_of(y0 + 1 + R));
        auto lv1 = fetch(xc, row_of(y0 - R));
        auto cd1 = centre(xc, row_of(y0 + 1));
        __syncthreads();   // pads zeroed before the first row is written
        int k = y0;
        for (; k + 1 < y1; k += 2) {
            step(k, en0, lv0, cd0);
            step(k + 1, en1, lv1, cd1);
        }
        if (k < y1) step(k, en0, lv0, cd0);
    } else {
        // Two rows per step (two barriers per row pair): rows k and k + 1
        // are published to two LDS row sets, their sums along x formed by
        // all threads at once; the next pair's loads are issued before the
        // first barrier.
        auto en0 = fetch(xc, row_of(y0 + R));
        auto lv0 = fetch(xc, row_of(y0 - R - 1));
        auto en1 = fetch(xc, row_of(y0 + 1 + R));
        auto lv1 = fetch(xc, row_of(y0 - R));
        auto cd0 = centre(xc, row_of(y0));
        auto cd1 = centre(xc, row_of(y0 + 1));
        __syncthreads();   // pads zeroed before the first row is written
        for (int k = y0; k < y1; k += 2) {
            slide(k, en0, lv0, rows);
            slide(k + 1, en1, lv1, rows + N * P);
            en0 = fetch(xc, row_of(k + 2 + R));
            lv0 = fetch(xc, row_of(k + 1 - R));
            en1 = fetch(xc, row_of(k + 3 + R));
            lv1 = fetch(xc, row_of(k + 2 - R));
            __syncthreads();
            runs();
            __syncthreads();
            finish(k, rows, cd0);
            if (k + 1 < y1) finish(k + 1, rows + N * P, cd1);
            cd0 = centre(xc, row_of(k + 2));
            cd1 = centre(xc, row_of(k + 3));
        }
    }
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

// The same two passes at a compile-time radius R (1..kRunMaxR) over
// strip_runs: the same per-pixel values and per-centre formulas, box sums
// formed as sliding sums instead of prefix differences; 1 / n is a constant
// for unclipped windows and a division only where the window is clipped.
constexpr int kRunMaxR = 16;

// Rows per strip_runs step, by pass (r = 8 on 6000 x 4000, µs, one / two
// rows per step): coefficients 325 / 277 (eight values: one row keeps half
// the threads idle in the sums along x), apply 548 / 558, backward centres
// 488 / 549 (two rows need 68 KB of LDS), backward pixels 601 / 610.
constexpr int kRpCoef = 2, kRpApply = 1, kRpCentres = 1, kRpPixels = 1;

template <int R> struct RunWindow {
    static constexpr int side = 2 * R + 1;
    static constexpr double inv_full = 1.0 / (double)(side * side);
};

template <int R>
__device__ __forceinline__ double inv_count(int cols, int rws) {
    if (cols == RunWindow<R>::side && rws == RunWindow<R>::side) return RunWindow<R>::inv_full;
    return 1.0 / (double)(cols * rws);
}

struct BackPixel {
    float p[3];
    float d[3];   // dL/dq at the pixel, clamped into the viewport (used only inside it)
};

// Branch-free: with an empty viewport the dL/dq load reads the pixel itself
// (its values are never used).
__device__ __forceinline__ BackPixel back_pixel(const GuidedBackArgs& g, int x, int y, bool have_view) {
    BackPixel o;
    const float* q = g.p + ((size_t)y * g.stride + (size_t)x) * 3;
    const int lx = max(min(x - g.vx, g.vw - 1), 0), ly = max(min(y - g.vy, g.vh - 1), 0);
    const float* d = have_view ? g.gq + ((size_t)ly * g.gstride + (size_t)lx) * 3 : q;
    o.p[0] = q[0]; o.p[1] = q[1]; o.p[2] = q[2];
    o.d[0] = d[0]; o.d[1] = d[1]; o.d[2] = d[2];
    return o;
}

template <int R>
__global__ void __launch_bounds__(kRunThreads, 3) gf_back_centres_runs(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double rows[kRpCentres * 14 * RunGeom<R>::plane];
    __shared__ double eps_part[kRunThreads / 64];
    const int span = kRunCols - 2 * R;
    const int strip_x = (int)blockIdx.x * span;
    const int y0 = (int)blockIdx.y * g.band, y1 = min(y0 + g.band, g.H);
    const bool have_view = g.vw > 0 && g.vh > 0;
    double geps = 0.0;
    auto fetch = [&](int x, int y) { return back_pixel(g, x, y, have_view); };
    auto eval = [&](const BackPixel& o, int x, int y, double v[14]) {
        const double pr = o.p[0], pg = o.p[1], pb = o.p[2];
        const double Y = (pg + g.wr * (pr - pg)) + g.wb * (pb - pg);
        v[0] = Y; v[1] = pr; v[2] = pg; v[3] = pb;
        v[4] = Y * Y; v[5] = Y * pr; v[6] = Y * pg; v[7] = Y * pb;
        const int lx = x - g.vx, ly = y - g.vy;
        if (lx >= 0 && ly >= 0 && lx < g.vw && ly < g.vh) {
            const double inv_n = inv_count<R>(min(g.W, x + R + 1) - max(0, x - R), min(g.H, y + R + 1) - max(0, y - R));
#pragma unroll
            for (int c = 0; c < 3; ++c) { v[8 + c] = (double)o.d[c] * Y * inv_n; v[11 + c] = (double)o.d[c] * inv_n; }
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) v[8 + c] = 0.0;
        }
    };
    auto centre = [&](int, int) { return 0; };
    const size_t plane = (size_t)g.W * g.H;
    auto sink = [&](int, int x, int y, const double S[14], int cols, int rws) {
        const double inv_n = inv_count<R>(cols, rws);
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
    strip_runs<14, R, kRpCentres>(g.W, g.H, 0, g.W, 0, g.H, strip_x, g.W, y0, y1, rows, fetch, eval, centre, sink);
    for (int off = 32; off > 0; off >>= 1) geps += __shfl_xor(geps, off, 64);
    if ((threadIdx.x & 63) == 0) eps_part[threadIdx.x >> 6] = geps;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kRunThreads / 64; ++w) t += eps_part[w];
        atomicAdd(g.geps, t);
    }
}

struct BackTerms {
    float w[11];
};

template <int R>
__global__ void __launch_bounds__(kRunThreads, 3) gf_back_pixels_runs(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double rows[kRpPixels * 11 * RunGeom<R>::plane];
    const int span = kRunCols - 2 * R;
    const int strip_x = (int)blockIdx.x * span;
    const int y0 = (int)blockIdx.y * g.band, y1 = min(y0 + g.band, g.H);
    const bool have_view = g.vw > 0 && g.vh > 0;
    const size_t plane = (size_t)g.W * g.H;
    auto fetch = [&](int x, int y) {
        BackTerms o;
        const float* w = g.ws + (size_t)y * g.W + (size_t)x;
#pragma unroll
        for (int m = 0; m < 11; ++m) o.w[m] = w[m * plane];
        return o;
    };
    auto eval = [&](const BackTerms& o, int, int, double v[11]) {
#pragma unroll
        for (int m = 0; m < 11; ++m) v[m] = (double)o.w[m];
    };
    auto centre = [&](int x, int y) { return back_pixel(g, x, y, have_view); };
    auto sink = [&](const BackPixel& o, int x, int y, const double S[11], int cols, int rws) {
        const double p[3] = {o.p[0], o.p[1], o.p[2]};
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
            const double inv_n = inv_count<R>(cols, rws);
#pragma unroll
            for (int c = 0; c < 3; ++c) gY += (double)o.d[c] * (S[8 + c] * inv_n);
        }
        gp[0] += gY * g.wr;
        gp[1] += gY * (1.0 - g.wr - g.wb);
        gp[2] += gY * g.wb;
        float* out = g.gp + ((size_t)y * g.W + (size_t)x) * 3;
        out[0] = (float)gp[0]; out[1] = (float)gp[1]; out[2] = (float)gp[2];
    };
    strip_runs<11, R, kRpPixels>(g.W, g.H, 0, g.W, 0, g.H, strip_x, g.W, y0, y1, rows, fetch, eval, centre, sink);
}

// The two-pass forward at a compile-time radius R (1..kRunMaxR) over
// strip_runs: gf_coef_runs writes a, b of each centre of the viewport grown
// by R as 6 FP32 planes, gf_apply_runs box-averages them and runs the
// post-filter program up to the first curve.  The same per-pixel moments and
// per-centre formulas as gf_coefficients / gf_apply.
// A pixel of the filtered level: at level 0 a plain load (branch-free, see
// strip_runs), at mip 1/2 the clipped native footprint's average.
template <bool Level0>
__device__ __forceinline__ void run_pixel(const GuidedArgs& g, int x, int y, float p[3]) {
    if (Level0) {
        const float* q = g.src + ((size_t)y * g.src_stride + (size_t)x) * 3;
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
    } else {
        level_pixel(g, x, y, p);
    }
}

template <int R, bool Level0>
__device__ __forceinline__ void coef_runs(const GuidedArgs& g, double* rows) {
    const int span = kRunCols - 2 * R;
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
    struct Pix { float p[3]; };
    auto fetch = [&](int x, int y) {
        Pix o;
        run_pixel<Level0>(g, x, y, o.p);
        return o;
    };
    auto eval = [&](const Pix& o, int, int, double m[8]) { moments(g, o.p, shift, m, err); };
    auto centre = [&](int, int) { return 0; };
    const size_t plane = (size_t)g.cw * g.ch;
    auto sink = [&](int, int kx, int k, const double S[8], int cols, int rws) {
        const double inv_n = inv_count<R>(cols, rws);
        const double my = S[0] * inv_n;
        double var = S[4] * inv_n - my * my;
        if (var < 0) var = 0;
        const double inv_d = 1.0 / (var + g.eps);
        float* c = g.coef + (size_t)(k - g.cy0) * g.cw + (size_t)(kx - g.cx0);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const double mp = S[1 + ch] * inv_n;
            const double cov = S[5 + ch] * inv_n - my * mp;
            const double a = cov * inv_d;
            // b in unshifted coordinates: mean(p) - a mean(Y).
            const double b = (mp - a * my) + (shift[ch] - a * shift_y);
            if (!__builtin_isfinite(a) || !__builtin_isfinite(b)) err |= FLAG_OVERFLOW;
            c[ch * plane] = (float)a;
            c[(3 + ch) * plane] = (float)b;
        }
    };
    // Only the pixels of the centres' windows enter the sums.
    const int c_lo = max(0, g.cx0 - R), c_hi = min(g.LW, g.cx0 + g.cw + R);
    strip_runs<8, R, kRpCoef>(g.LW, g.LH, c_lo, c_hi, 0, g.LH, strip_x, g.cx0 + g.cw, k0, k1, rows, fetch, eval, centre,
                     sink);
    if (err && g.flags) atomicOr(g.flags, err);
}

template <int R>
__global__ void __launch_bounds__(kRunThreads, 3) gf_coef_runs(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    __shared__ double rows[kRpCoef * 8 * RunGeom<R>::plane];
    if (g.mip == 0)
        coef_runs<R, true>(g, rows);
    else
        coef_runs<R, false>(g, rows);
}

template <int R, bool Level0>
__device__ __forceinline__ void apply_runs(const GuidedArgs& g, double* rows, double* tables) {
    const int span = kRunCols - 2 * R;
    const int strip_x = g.vx + (int)blockIdx.x * span;
    const int i0 = g.vy + (int)blockIdx.y * g.band;
    const int i1 = min(i0 + g.band, g.vy + g.vh);
    uint32_t err = 0;
    // Coefficients exist for centres in [cx0, cx0 + cw) x [cy0, cy0 + ch):
    // exactly the clipped windows of the viewport's pixels.
    const size_t plane = (size_t)g.cw * g.ch;
    struct Coef { float c[6]; };
    auto fetch = [&](int kx, int ky) {
        Coef o;
        const float* q = g.coef + (size_t)(ky - g.cy0) * g.cw + (size_t)(kx - g.cx0);
#pragma unroll
        for (int m = 0; m < 6; ++m) o.c[m] = q[m * plane];
        return o;
    };
    auto eval = [&](const Coef& o, int, int, double c[6]) {
#pragma unroll
        for (int m = 0; m < 6; ++m) c[m] = (double)o.c[m];
    };
    struct Pix { float p[3]; };
    auto centre = [&](int x, int i) {
        Pix o;
        run_pixel<Level0>(g, x, i, o.p);
        return o;
    };
    // The post-filter ops from g.fused on (g.fused is a curve op) run here
    // too, the curve op's knot tables staged in LDS when every curve has at
    // most kTailKnots knots.
    Const<OpList>& L = as_const(*g.post);
    const uint32_t end = L.n;
    bool staged = false;
    LdsPoly curves[4];
    uint32_t cubic = 0;
    if (g.fused < end) {
        const DevCurves* cv = L.op[g.fused].curves;
        staged = true;
        for (int c = 0; c < 4; ++c) staged = staged && (as_const(*cv).c[c].identity || as_const(*cv).c[c].n <= kTailKnots);
        cubic = as_const(*cv).cubic;
        if (staged)
            for (int c = 0; c < 4; ++c) curves[c] = stage_poly(&cv->c[c], cubic, tables + c * poly_table_doubles(kTailKnots));
    }
    auto sink = [&](const Pix& px, int x, int i, const double S[6], int cols, int rws) {
        const double inv_n = inv_count<R>(cols, rws);
        const float* p = px.p;
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
        if (staged) {
            curves3_lds(curves, o, err);
            apply_ops(*g.post, g.fused + 1, end, o[0], o[1], o[2], err);
        } else {
            apply_ops(*g.post, g.fused, end, o[0], o[1], o[2], err);
        }
        float* d = g.out + ((size_t)(i - g.vy) * g.ostride + (size_t)(x - g.vx)) * 3;
        d[0] = o[0]; d[1] = o[1]; d[2] = o[2];
    };
    strip_runs<6, R, kRpApply>(g.LW, g.LH, g.cx0, g.cx0 + g.cw, g.cy0, g.cy0 + g.ch, strip_x, g.vx + g.vw, i0, i1, rows, fetch,
                     eval, centre, sink);
    if (err && g.flags) atomicOr(g.flags, err);
}

template <int R>
__global__ void __launch_bounds__(kRunThreads, 3) gf_apply_runs(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    __shared__ double rows[kRpApply * 6 * RunGeom<R>::plane];
    __shared__ __attribute__((aligned(16))) double tables[4 * poly_table_doubles(kTailKnots)];
    if (g.mip == 0)
        apply_runs<R, true>(g, rows, tables);
    else
        apply_runs<R, false>(g, rows, tables);
}

// Static LDS at the largest radius: four workgroups of either pass fit one CU.
static_assert(sizeof(double) * (kRpCentres * 14 * RunGeom<kRunMaxR>::plane + 4) <= 160 * 1024 / 4, "gf_back_centres_runs LDS");

// Rows per band for the strip_runs passes: about three workgroups per CU
// over one image, at least 32 rows (each band repeats the 2R + 1 row
// warm-up; stacks of 512^2 crops run 1.6x faster than with 64-row bands).
// Depends on the image only, so a stack equals single calls bit for bit.
inline int run_band_rows(int strips, int rows) {
    const int per = std::max(1, 768 / std::max(1, strips));
    return std::max(32, (rows + per - 1) / per);
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
    if (g.r >= 1 && g.r <= kRunMaxR) {
        const int span = kRunCols - 2 * g.r;
        const int strips = (g.W + span - 1) / span;
        g.band = run_band_rows(strips, g.H);
        dim3 grid((unsigned)strips, (unsigned)((g.H + g.band - 1) / g.band), nz);
        switch (g.r) {
#define RAWOPS_GF_BACK(R) \
    case R: \
        hipLaunchKernelGGL(gf_back_centres_runs<R>, grid, dim3(kRunThreads), 0, stream, g); \
        if (hipGetLastError() != hipSuccess) break; \
        hipLaunchKernelGGL(gf_back_pixels_runs<R>, grid, dim3(kRunThreads), 0, stream, g); \
        break;
            RAWOPS_GF_BACK(1) RAWOPS_GF_BACK(2) RAWOPS_GF_BACK(3) RAWOPS_GF_BACK(4) RAWOPS_GF_BACK(5)
            RAWOPS_GF_BACK(6) RAWOPS_GF_BACK(7) RAWOPS_GF_BACK(8) RAWOPS_GF_BACK(9) RAWOPS_GF_BACK(10)
            RAWOPS_GF_BACK(11) RAWOPS_GF_BACK(12) RAWOPS_GF_BACK(13) RAWOPS_GF_BACK(14) RAWOPS_GF_BACK(15)
            RAWOPS_GF_BACK(16)
#undef RAWOPS_GF_BACK
        }
        return hipGetLastError();
    }
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
    g.fused = L.fused_ops;
    if (g.r <= kRunMaxR) {
        // Two strip passes over sliding sums; bands depend on the image only.
        const int span = kRunCols - 2 * g.r;
        const int strips_c = (g.cw + span - 1) / span;
        GuidedArgs gc = g;
        gc.band = run_band_rows(strips_c, g.ch);
        dim3 grid_c((unsigned)strips_c, (unsigned)((g.ch + gc.band - 1) / gc.band), nz);
        const int strips_a = (g.vw + span - 1) / span;
        GuidedArgs ga = g;
        ga.band = run_band_rows(strips_a, g.vh);
        dim3 grid_a((unsigned)strips_a, (unsigned)((g.vh + ga.band - 1) / ga.band), nz);
        hipError_t e = hipSuccess;
        switch (g.r) {
#define RAWOPS_GF_RUNS(R) \
    case R: \
        hipLaunchKernelGGL(gf_coef_runs<R>, grid_c, dim3(kRunThreads), 0, stream, gc); \
        if ((e = hipGetLastError()) != hipSuccess) break; \
        hipLaunchKernelGGL(gf_apply_runs<R>, grid_a, dim3(kRunThreads), 0, stream, ga); \
        e = hipGetLastError(); \
        break;
            RAWOPS_GF_RUNS(1) RAWOPS_GF_RUNS(2) RAWOPS_GF_RUNS(3) RAWOPS_GF_RUNS(4) RAWOPS_GF_RUNS(5)
            RAWOPS_GF_RUNS(6) RAWOPS_GF_RUNS(7) RAWOPS_GF_RUNS(8) RAWOPS_GF_RUNS(9) RAWOPS_GF_RUNS(10)
            RAWOPS_GF_RUNS(11) RAWOPS_GF_RUNS(12) RAWOPS_GF_RUNS(13) RAWOPS_GF_RUNS(14) RAWOPS_GF_RUNS(15)
            RAWOPS_GF_RUNS(16)
#undef RAWOPS_GF_RUNS
        }
        return e;
    }
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
