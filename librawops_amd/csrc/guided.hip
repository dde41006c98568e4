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
    __shared__ __attribute__((aligned(16))) double tables[4 * (kPolyRec + 1) * kTailKnots];
    const size_t n = (size_t)g.vw * g.vh;
    Const<OpList>& L = as_const(*g.post);
    const uint32_t end = L.n;
    const DevCurves* cv = L.op[g.fused].curves;
    bool staged = true;
    for (int c = 0; c < 4; ++c) staged = staged && (as_const(*cv).c[c].identity || as_const(*cv).c[c].n <= kTailKnots);
    LdsPoly curves[4];
    const uint32_t cubic = as_const(*cv).cubic;
    if (staged) {
        for (int c = 0; c < 4; ++c) curves[c] = stage_poly(&cv->c[c], tables + c * (kPolyRec + 1) * kTailKnots);
        __syncthreads();
    }
    uint32_t err = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / g.vw), x = (int)(i - (size_t)y * g.vw);
        float* d = g.out + ((size_t)y * g.ostride + (size_t)x) * 3;
        float p[3] = {d[0], d[1], d[2]};
        if (staged) {
            curves3_lds(curves, cubic, p, err);
            apply_ops(*g.post, g.fused + 1, end, p[0], p[1], p[2], err);
        } else {
            apply_ops(*g.post, g.fused, end, p[0], p[1], p[2], err);
        }
        d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
    }
    if (err && g.flags) atomicOr(g.flags, err);
}

// ---------------------------------------------------------------- fused --
//
// The whole filter and its post-filter program in one pass at level 0, with
// the per-centre coefficients kept on chip instead of a workspace round trip.
// A workgroup (8 wavefronts) owns a strip of kFusedCols pixel columns
// j = 0..255 at x = px0 + j, px0 = ox - 2r, and a band of output rows; the
// centres of the strip are j in [r, 256 - r), its outputs j in [2r, 256 - 2r)
// (Wo = 256 - 4r columns from x = ox).  Step s, with c = c_lo + s:
//
//   wave 0: the vertical window sums of Y and Y^2 of centre row c, per pixel
//     column (a lane holds 4 adjacent columns), slid down one row per step;
//     their sums along x as sliding sums over the lane's run of 4 centres,
//     read back from a wave-private LDS row zero-padded by r at both ends;
//     each centre's mean Y and 1 / (var + eps), written to a double-buffered
//     row.
//   waves 1..3 (channel ch): the same for p_ch and Y p_ch of row c (kept to
//     the next step), and, from last step's sums and wave 0's row, a_ch and
//     b_ch of centre row c - 1, summed along x over each output column's
//     r-window the same way and written (FP32) into a ring of 2r + 3 centre
//     rows in LDS.
//   waves 4..7: a thread per output column adds the ring row c - 2 to the
//     sums of its window and subtracts row c - 2r - 3, forms output row
//     c - 2 - r and runs the post-filter program (curves included), and
//     stores the pixel.
//
// One barrier per step; the ring slot written in a step is the one read as
// the leaving row a step later.  One workgroup per CU (ring of 2r + 3 rows x
// 6 planes; 98 KB at r = 8).
constexpr int kFusedCols = 256;
constexpr int kFusedThreads = 512;
constexpr int kFusedMaxR = 11;
constexpr int kFusedRun = 4;                                        // columns per coefficient lane
constexpr int kFusedLogical = kFusedCols + 2 * kFusedMaxR;          // zero-padded run row
constexpr int kFusedPlane = kFusedLogical + (kFusedLogical >> 2);   // + one pad double per 4 (banks)
constexpr int kFusedMinBand = 64;                                   // rows per band at least

// Ring geometry at radius R: a row of each plane holds the outputs from
// column RB = 2R rounded down to a multiple of 4 (so a coefficient lane's 4
// outputs are one aligned 16-byte group), WR floats (a multiple of 4).
template <int R> struct FusedRing {
    static constexpr int D = 2 * R + 3;
    static constexpr int RB = (2 * R) & ~3;
    static constexpr int WR = ((kFusedCols - 2 * R - RB) + 3) & ~3;
    static constexpr int floats = D * 6 * WR;
};

__device__ __forceinline__ int run_index(int i) { return i + (i >> 2); }

__device__ __forceinline__ int clipped_count(int v, int r, int n) {
    return min(n, v + r + 1) - max(0, v - r);
}

// Four adjacent pixels of row yy at columns x0..x0+3; columns outside
// [0, W) load a clamped in-bounds pixel (its moments are discarded).
__device__ __forceinline__ void load4(const float* __restrict__ src, uint32_t stride, int yy, int x0, int W,
                                      float p[12]) {
    const float* row = src + (size_t)yy * stride * 3;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int x = min(max(x0 + t, 0), W - 1);
        const float* q = row + (size_t)x * 3;
        p[3 * t] = q[0]; p[3 * t + 1] = q[1]; p[3 * t + 2] = q[2];
    }
}

// Sliding sums over a zero-padded run row: s[t] = sum_{i=0}^{2R} b[i0 + t + i]
// (two partial sums for the first window, so the adds overlap).
template <int R>
__device__ __forceinline__ void run_sums(const double* __restrict__ b, int i0, double s[4]) {
    double e = 0.0, o = 0.0;
#pragma unroll
    for (int i = 0; i <= 2 * R; i += 2) e += b[run_index(i0 + i)];
#pragma unroll
    for (int i = 1; i <= 2 * R; i += 2) o += b[run_index(i0 + i)];
    double acc = e + o;
    s[0] = acc;
#pragma unroll
    for (int t = 1; t < 4; ++t) {
        acc = acc + b[run_index(i0 + 2 * R + t)] - b[run_index(i0 + t - 1)];
        s[t] = acc;
    }
}

// Box sums for a compile-time radius R (1..11) without per-row prefix scans.
// The strip is kRunCols = 256 pixel columns x = strip_x - R + t; thread t
// keeps the vertical window sums (rows k-R..k+R, clipped) of the N values of
// its column, slid down one row per step, with the row data of the next step
// loaded a step ahead (fetch(x, y) returns it, eval(data, x, y, v) turns it
// into the N values).  Per row:
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
// to [0, LW) x [0, LH); only columns in [c_lo, c_hi) enter the sums; fetch
// and centre are called at clamped in-bounds coordinates, so their loads are
// unconditional.
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

// s[t] = sum_{i=0}^{2R} b[run_slot(i0 + t + i)], t < kRunLen, as sliding sums
// (two partial sums for the first window, so the adds overlap).
template <int R>
__device__ __forceinline__ void run_sums16(const double* __restrict__ b, int i0, double s[kRunLen]) {
    double e = 0.0, o = 0.0;
#pragma unroll
    for (int i = 0; i <= 2 * R; i += 2) e += b[run_slot(i0 + i)];
#pragma unroll
    for (int i = 1; i <= 2 * R; i += 2) o += b[run_slot(i0 + i)];
    double acc = e + o;
    s[0] = acc;
#pragma unroll
    for (int t = 1; t < kRunLen; ++t) {
        acc = acc + b[run_slot(i0 + 2 * R + t)] - b[run_slot(i0 + t - 1)];
        s[t] = acc;
    }
}

template <int N, int R, class Fetch, class Eval, class Centre, class Sink>
__device__ __forceinline__ void strip_runs(int LW, int LH, int c_lo, int c_hi, int strip_x, int xend, int y0, int y1,
                                           double* rows /* [N][plane] */, Fetch fetch, Eval eval, Centre centre,
                                           Sink sink) {
    static_assert(N * (kRunCols / kRunLen) <= kRunThreads, "one sum run per thread");
    constexpr int P = RunGeom<R>::plane;
    const int t = threadIdx.x;
    const int x = strip_x - R + t;
    const bool own = x >= c_lo && x < c_hi;
    const int xc = min(max(x, 0), LW - 1);
    const int slot = run_slot(t + R);
    const int vm = t / (kRunCols / kRunLen), vj = t % (kRunCols / kRunLen);   // value, run of the sums along x
    auto row_of = [&](int y) { return min(max(y, 0), LH - 1); };
    for (int i = t; i < N * P; i += kRunThreads) rows[i] = 0.0;
    double V[N];
#pragma unroll
    for (int m = 0; m < N; ++m) V[m] = 0.0;
    if (own) {
        const int ya = max(0, y0 - R), yb = min(LH, y0 + R + 1);
        for (int y = ya; y < yb; ++y) {
            const auto d = fetch(xc, y);
            double v[N];
            eval(d, x, y, v);
#pragma unroll
            for (int m = 0; m < N; ++m) V[m] += v[m];
        }
    }
    auto en = fetch(xc, row_of(y0 + 1 + R));
    auto lv = fetch(xc, row_of(y0 - R));
    const bool is_centre = t >= R && t < kRunCols - R && x < xend;
    auto cd = centre(xc, row_of(y0));
    __syncthreads();   // pads zeroed before the first row is written
    for (int k = y0; k < y1; ++k) {
        if (k > y0) {
            if (own) {
                double v[N];
                if (k + R < LH) {
                    eval(en, x, k + R, v);
#pragma unroll
                    for (int m = 0; m < N; ++m) V[m] += v[m];
                }
                if (k - R - 1 >= 0) {
                    eval(lv, x, k - R - 1, v);
#pragma unroll
                    for (int m = 0; m < N; ++m) V[m] -= v[m];
                }
            }
            en = fetch(xc, row_of(k + 1 + R));
            lv = fetch(xc, row_of(k - R));
        }
#pragma unroll
        for (int m = 0; m < N; ++m) rows[m * P + slot] = V[m];
        __syncthreads();
        if (vm < N) {
            double* b = rows + vm * P;
            double Hs[kRunLen];
            run_sums16<R>(b, kRunLen * vj, Hs);
            // Every lane's reads of the wavefront's rows are done before any
            // lane overwrites them.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int u = 0; u < kRunLen; ++u) b[run_slot(kRunLen * vj + u + R)] = Hs[u];
        }
        __syncthreads();
        const auto cur = cd;
        if (k + 1 < y1) cd = centre(xc, row_of(k + 1));
        if (is_centre) {
            double S[N];
#pragma unroll
            for (int m = 0; m < N; ++m) S[m] = rows[m * P + slot];
            const int cols = min(LW, x + R + 1) - max(0, x - R);
            const int rws = min(LH, k + R + 1) - max(0, k - R);
            sink(cur, x, k, S, cols, rws);
        }
    }
}

template <int R>
__global__ void __launch_bounds__(kFusedThreads) gf_fused(const GuidedArgs g0) {
    const GuidedArgs g = slice_of(g0);
    using RG = FusedRing<R>;
    constexpr int D = RG::D, RB = RG::RB, WR = RG::WR;
    __shared__ __attribute__((aligned(16))) float ring[RG::floats];
    __shared__ double runs[4][2][kFusedPlane];
    __shared__ double my_d[2][2][4][64];        // [buffer][mean Y, 1/(var+eps)][column of the lane][lane]
    const int ox = g.vx + (int)blockIdx.x * (kFusedCols - 4 * R), px0 = ox - 2 * R;
    const int y0 = g.vy + (int)blockIdx.y * g.band, y1 = min(y0 + g.band, g.vy + g.vh);
    const int c_lo = max(0, y0 - R), c_hi = min(g.LH, y1 + R);
    const int steps = y1 + R - c_lo + 2;
    // Pixel columns inside the support of the viewport (its 2R halo): only
    // these enter the moments and the non-finite check.
    const int sx0 = max(0, g.vx - 2 * R), sx1 = min(g.LW, g.vx + g.vw + 2 * R);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    uint32_t err = 0;

    // Zero the run rows (their pads stay zero).
    for (int i = t; i < 4 * 2 * kFusedPlane; i += kFusedThreads) (&runs[0][0][0])[i] = 0.0;
    // Per-strip shift of the moments: the strip's first centre pixel.
    double shift[3];
    {
        const float* q = g.src + ((size_t)c_lo * g.src_stride + (size_t)min(max(px0 + R, sx0), sx1 - 1)) * 3;
        shift[0] = q[0]; shift[1] = q[1]; shift[2] = q[2];
    }
    const double shift_y = working_y(g, shift[0], shift[1], shift[2]);
    __syncthreads();

    if (wave < 4) {
        // ---- coefficient waves
        const int role = wave;                 // 0: Y, Y^2; 1 + ch: p_ch, Y p_ch
        const int ch = role - 1;
        const int j0 = kFusedRun * lane;       // first pixel column of the lane
        const int x0 = px0 + j0;
        // colok: the column enters the moments; centre: a centre of the strip
        // inside the image (others are zero in the sums along x); need: a
        // centre some output of the viewport reads (checked for overflow).
        bool colok[4], centre[4], need[4];
        double nxc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            colok[u] = x0 + u >= sx0 && x0 + u < sx1;
            centre[u] = j0 + u >= R && j0 + u < kFusedCols - R && x0 + u >= 0 && x0 + u < g.LW;
            need[u] = centre[u] && x0 + u >= g.vx - R && x0 + u < g.vx + g.vw + R;
            nxc[u] = (double)clipped_count(x0 + u, R, g.LW);
        }
        double V0[4] = {0, 0, 0, 0}, V1[4] = {0, 0, 0, 0};
        auto accumulate = [&](const float p[12], double sign, bool check) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float* q = p + 3 * u;
                if (check && colok[u] && (!finite(q[0]) || !finite(q[1]) || !finite(q[2]))) err |= FLAG_INVALID;
                const double d0 = (double)q[0] - shift[0], d1 = (double)q[1] - shift[1], d2 = (double)q[2] - shift[2];
                const double y = working_y(g, d0, d1, d2);
                const double dc = ch == 0 ? d0 : (ch == 1 ? d1 : d2);
                const double m0 = role == 0 ? y : dc;
                const double m1 = role == 0 ? y * y : y * dc;
                V0[u] = colok[u] ? V0[u] + sign * m0 : V0[u];
                V1[u] = colok[u] ? V1[u] + sign * m1 : V1[u];
            }
        };
        // Window of centre row c_lo.
        for (int yy = max(0, c_lo - R); yy <= min(g.LH - 1, c_lo + R); ++yy) {
            float p[12];
            load4(g.src, g.src_stride, yy, x0, g.LW, p);
            accumulate(p, 1.0, role == 0);
        }
        double* b0 = runs[wave][0];
        double* b1 = runs[wave][1];
        // Rows of step s arrive two steps ahead (buffer s & 1), so a load has
        // two steps to return before its moments are formed.
        // The loads are unconditional (rows clamped into the image; unused
        // rows are never accumulated), so the compiler's count of loads in
        // flight stays exact and a wait covers only the rows it needs.
        float enter[2][12], leave[2][12];
        auto fetch = [&](int c, int buf) {
            load4(g.src, g.src_stride, min(c + R, g.LH - 1), x0, g.LW, enter[buf]);
            load4(g.src, g.src_stride, max(c - R - 1, 0), x0, g.LW, leave[buf]);
        };
        const double sh = ch == 0 ? shift[0] : (ch == 1 ? shift[1] : shift[2]);
        double P0[4], P1[4], Pn[4];            // channel waves: last step's sums and 1/n
        auto step = [&](int s, int buf) {
            const int c = c_lo + s;
            if (role > 0 && s >= 1 && c - 1 < c_hi) {
                // a, b of centre row c - 1 and their sums along x -> ring
                const double* md = &my_d[(s - 1) & 1][0][0][0];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + u;
                    const double my = md[u * 64 + lane], inv_d = md[256 + u * 64 + lane];
                    const double mp = P0[u] * Pn[u];
                    const double cov = P1[u] * Pn[u] - my * mp;
                    const double a = cov * inv_d;
                    const double b = (mp - a * my) + (sh - a * shift_y);
                    if (need[u] && (!__builtin_isfinite(a) || !__builtin_isfinite(b))) err |= FLAG_OVERFLOW;
                    b0[run_index(j + R)] = centre[u] ? a : 0.0;
                    b1[run_index(j + R)] = centre[u] ? b : 0.0;
                }
                double Ha[4], Hb[4];
                run_sums<R>(b0, j0, Ha);
                run_sums<R>(b1, j0, Hb);
                float* slot = ring + ((c - 1) % D) * 6 * WR;
                if (j0 >= 2 * R && j0 + 3 < kFusedCols - 2 * R) {
                    *reinterpret_cast<float4*>(slot + ch * WR + j0 - RB) =
                        make_float4((float)Ha[0], (float)Ha[1], (float)Ha[2], (float)Ha[3]);
                    *reinterpret_cast<float4*>(slot + (3 + ch) * WR + j0 - RB) =
                        make_float4((float)Hb[0], (float)Hb[1], (float)Hb[2], (float)Hb[3]);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int jo = j0 + u;
                        if (jo >= 2 * R && jo < kFusedCols - 2 * R) {
                            slot[ch * WR + jo - RB] = (float)Ha[u];
                            slot[(3 + ch) * WR + jo - RB] = (float)Hb[u];
                        }
                    }
                }
            }
            if (c < c_hi) {
                // moments of centre row c
                if (s > 0 && c + R < g.LH) accumulate(enter[buf], 1.0, role == 0);
                if (s > 0 && c - R - 1 >= 0) accumulate(leave[buf], -1.0, false);
            }
            fetch(min(c + 2, c_hi - 1), buf);
            if (c < c_hi) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    b0[run_index(j0 + u + R)] = V0[u];
                    b1[run_index(j0 + u + R)] = V1[u];
                }
                double S0[4], S1[4];
                run_sums<R>(b0, j0, S0);
                run_sums<R>(b1, j0, S1);
                // 1 / n of each centre: one division per row, another only
                // where the window is clipped at the left or right edge.
                const int ny = clipped_count(c, R, g.LH);
                const double inv_row = 1.0 / (double)((2 * R + 1) * ny);
                double inv_n[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    inv_n[u] = inv_row;
                    if (nxc[u] != (double)(2 * R + 1)) inv_n[u] = 1.0 / (nxc[u] * (double)ny);
                }
                if (role == 0) {
                    double* md = &my_d[s & 1][0][0][0];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double my = S0[u] * inv_n[u];
                        double var = S1[u] * inv_n[u] - my * my;
                        if (var < 0) var = 0;
                        md[u * 64 + lane] = my;
                        md[256 + u * 64 + lane] = 1.0 / (var + g.eps);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) { P0[u] = S0[u]; P1[u] = S1[u]; Pn[u] = inv_n[u]; }
                }
            }
            __syncthreads();
        };
        fetch(c_lo + 1, 1);
        for (int s = 0; s < steps; s += 2) {
            step(s, 0);
            if (s + 1 < steps) step(s + 1, 1);
        }
    } else {
        // ---- output waves
        const int jo = t - 256;
        const int x = px0 + jo;
        const bool col = jo >= 2 * R && jo < kFusedCols - 2 * R && x < g.vx + g.vw;
        const int ri = min(max(jo - RB, 0), WR - 1);
        const int nx = clipped_count(x, R, g.LW);
        double U[6] = {0, 0, 0, 0, 0, 0};
        float pix[3] = {0, 0, 0};
        // Unconditional load at a clamped in-bounds pixel (see above).
        const int xc = min(max(x, 0), g.LW - 1);
        auto fetch = [&](int k) {
            const float* q = g.src + ((size_t)min(max(k, y0), y1 - 1) * g.src_stride + (size_t)xc) * 3;
            pix[0] = q[0]; pix[1] = q[1]; pix[2] = q[2];
        };
        fetch(c_lo - 2 - R);
        for (int s = 0; s < steps; ++s) {
            const int cb = c_lo + s - 2;       // centre row written into the ring last step
            const int k = cb - R;              // output row
            if (s >= 2) {
                if (cb < c_hi) {
                    const float* a = ring + (cb % D) * 6 * WR + ri;
#pragma unroll
                    for (int p = 0; p < 6; ++p) U[p] += (double)a[p * WR];
                }
                if (cb - 2 * R - 1 >= c_lo) {
                    const float* a = ring + ((cb - 2 * R - 1) % D) * 6 * WR + ri;
#pragma unroll
                    for (int p = 0; p < 6; ++p) U[p] -= (double)a[p * WR];
                }
            }
            const float p0 = pix[0], p1 = pix[1], p2 = pix[2];
            fetch(k + 1);
            if (s >= 2 && k >= y0 && k < y1) {
                // 1 / n: one division per row, another only at the left or right edge.
                const int ny = clipped_count(k, R, g.LH);
                double inv_n = 1.0 / (double)((2 * R + 1) * ny);
                if (nx != 2 * R + 1) inv_n = 1.0 / (double)(nx * ny);
                if (col) {
                    const float p[3] = {p0, p1, p2};
                    const double y = working_y(g, (double)p0, (double)p1, (double)p2);
                    float o[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double q = (U[c] * inv_n) * y + U[3 + c] * inv_n;
                        const double delta = q - (double)p[c];
                        if (!__builtin_isfinite(delta)) err |= FLAG_OVERFLOW;
                        if (delta == 0) { o[c] = p[c]; continue; }   // exact copy, as the reference
                        const double value = (double)p[c] + delta;
                        if (value < -3.4028234663852886e38 || value > 3.4028234663852886e38) err |= FLAG_OVERFLOW;
                        o[c] = (float)value;
                    }
                    apply_ops(*g.post, 0, g.fused, o[0], o[1], o[2], err);
                    float* d = g.out + ((size_t)(k - g.vy) * g.ostride + (size_t)(x - g.vx)) * 3;
                    d[0] = o[0]; d[1] = o[1]; d[2] = o[2];
                }
            }
            __syncthreads();
        }
    }
    if (err && g.flags) atomicOr(g.flags, err);
}

// Static LDS of gf_fused at its largest radius.
static_assert(sizeof(float) * FusedRing<kFusedMaxR>::floats + sizeof(double) * (4 * 2 * kFusedPlane + 2 * 2 * 4 * 64) <=
                  160 * 1024,
              "gf_fused LDS");

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
constexpr int kRunMaxR = 11;

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

__device__ __forceinline__ BackPixel back_pixel(const GuidedBackArgs& g, int x, int y, bool have_view) {
    BackPixel o;
    const float* q = g.p + ((size_t)y * g.stride + (size_t)x) * 3;
    o.p[0] = q[0]; o.p[1] = q[1]; o.p[2] = q[2];
    if (have_view) {
        const int lx = min(max(x - g.vx, 0), g.vw - 1), ly = min(max(y - g.vy, 0), g.vh - 1);
        const float* d = g.gq + ((size_t)ly * g.gstride + (size_t)lx) * 3;
        o.d[0] = d[0]; o.d[1] = d[1]; o.d[2] = d[2];
    } else {
        o.d[0] = o.d[1] = o.d[2] = 0.0f;
    }
    return o;
}

template <int R>
__global__ void __launch_bounds__(kRunThreads, 3) gf_back_centres_runs(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double rows[14 * RunGeom<R>::plane];
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
    strip_runs<14, R>(g.W, g.H, 0, g.W, strip_x, g.W, y0, y1, rows, fetch, eval, centre, sink);
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
__global__ void __launch_bounds__(kRunThreads, 4) gf_back_pixels_runs(const GuidedBackArgs g0) {
    const GuidedBackArgs g = slice_of(g0);
    __shared__ double rows[11 * RunGeom<R>::plane];
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
    strip_runs<11, R>(g.W, g.H, 0, g.W, strip_x, g.W, y0, y1, rows, fetch, eval, centre, sink);
}

// Static LDS at the largest radius: four workgroups of either pass fit one CU.
static_assert(sizeof(double) * (14 * RunGeom<kRunMaxR>::plane + 4) <= 160 * 1024 / 4, "gf_back_centres_runs LDS");

// Rows per band for the strip_runs passes: about three workgroups per CU
// over one image, at least 64 rows (each band repeats the 2R + 1 row
// warm-up).  Depends on the image only, so a stack equals single calls bit
// for bit.
inline int run_band_rows(int strips, int rows) {
    const int per = std::max(1, 768 / std::max(1, strips));
    return std::max(64, (rows + per - 1) / per);
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
            RAWOPS_GF_BACK(11)
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
    if (g.mip == 0 && g.r <= kFusedMaxR) {
        // One pass, coefficients on chip.  Bands: about one workgroup per CU
        // over the image (at least kFusedMinBand rows, which amortise the
        // 2r + 1 steps before a band's first output); the band height
        // depends on the image only, so a stack filters each image as a
        // single call does, bit for bit.
        const int Wo = kFusedCols - 4 * g.r;
        const int strips = (g.vw + Wo - 1) / Wo;
        const int per = std::max(1, 256 / strips);
        g.band = std::max(kFusedMinBand, (g.vh + per - 1) / per);
        g.fused = L.fused_ops;
        dim3 grid((unsigned)strips, (unsigned)((g.vh + g.band - 1) / g.band), nz);
        switch (g.r) {
#define RAWOPS_GF_FUSED(R) case R: hipLaunchKernelGGL(gf_fused<R>, grid, dim3(kFusedThreads), 0, stream, g); break;
            RAWOPS_GF_FUSED(1) RAWOPS_GF_FUSED(2) RAWOPS_GF_FUSED(3) RAWOPS_GF_FUSED(4) RAWOPS_GF_FUSED(5)
            RAWOPS_GF_FUSED(6) RAWOPS_GF_FUSED(7) RAWOPS_GF_FUSED(8) RAWOPS_GF_FUSED(9) RAWOPS_GF_FUSED(10)
            RAWOPS_GF_FUSED(11)
#undef RAWOPS_GF_FUSED
        }
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && L.fused_ops < L.post_ops) {
            const size_t n = (size_t)g.vw * g.vh;
            const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535);
            hipLaunchKernelGGL(gf_ops_tail, dim3(blocks, 1, nz), dim3(256), 0, stream, g);
            e = hipGetLastError();
        }
        return e;
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
