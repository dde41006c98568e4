// Streaming Menon (DDFAPD) render of the interior of a viewport: register
// windows and wavefront lane shifts instead of LDS tiles.
//
// A wavefront owns a strip of 128 sensor columns, two per lane (lane L holds
// columns xA = X0 + 112 w - 8 + 2 L and xA + 1), and walks down a band of
// rows.  Every Menon stage (MenonDemosaic.cpp:89-174) runs once per row on
// the lane's one non-green and one green site, in the reference's FP64
// operation order (this file is compiled with -ffp-contract=off), and keeps
// the rows later stages read in a sliding register window:
//
//   step s (final row s):  raw row s+6 arrives (loaded PF rows ahead)
//                          directional greens, ch and dh at row s+4
//                          cv and dv at row s+2, classifier at row s+2
//                          red / blue at the green site of row s+1
//                          final RGB of row s, the op program, stores
//
// Horizontal neighbours at x +- 1 and x +- 2 are the lane's own other column
// or a column of lane L -+ 1, read with DPP wave_shr:1 / wave_shl:1 (lane L
// receives lane L - 1's / L + 1's value); vertical neighbours are rows of
// the lane's own window.  The support of an output pixel is +-6 columns, so
// lanes 4..59 (112 columns) hold valid results; lanes 0..3 and 60..63 are the
// strip's halo.  Strips that reach the edge of the active area run the edge
// form (reflected loads, zero gradients outside the area, mirrored rows and
// columns; see Strip), the others a form with no edge logic.
#include <algorithm>

#include "render_launch.hpp"
#include "stream_launch.hpp"

namespace rawops {
namespace {

constexpr int kStreamThreads = 256;   // four wavefronts, side by side
constexpr int kPrefetch = 4;          // raw rows in flight ahead of use

// bound_ctrl: lane 0 (wave_shr) / lane 63 (wave_shl) read 0, so no "old"
// value has to be set up in the destination first.
__device__ __forceinline__ int lane_prev(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int lane_next(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ float prev_f(float v) { return __builtin_bit_cast(float, lane_prev(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ float next_f(float v) { return __builtin_bit_cast(float, lane_next(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ double prev_d(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)lane_prev((int)(uint32_t)b), hi = (uint32_t)lane_prev((int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double next_d(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)lane_next((int)(uint32_t)b), hi = (uint32_t)lane_next((int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The same moves for the element type of the windows (FP64, or FP32 in the
// approximate form).
__device__ __forceinline__ float prev_t(float v) { return prev_f(v); }
__device__ __forceinline__ double prev_t(double v) { return prev_d(v); }
__device__ __forceinline__ float next_t(float v) { return next_f(v); }
__device__ __forceinline__ double next_t(double v) { return next_d(v); }
__device__ __forceinline__ float abs_t(float v) { return __builtin_fabsf(v); }
__device__ __forceinline__ double abs_t(double v) { return __builtin_fabs(v); }

// Whole-sample reflection into [lo, hi] (MenonDemosaic.cpp:60-66) for a
// coordinate at most one area size outside, then clamped so that far-out
// positions (whose values no output reads) still load in bounds.
__device__ __forceinline__ int reflect_in(int v, int lo, int hi) {
    v = v < lo ? 2 * lo - v : (v > hi ? 2 * hi - v : v);
    return min(max(v, lo), hi);
}

// Value of lane `src` (any lane of the wavefront).
__device__ __forceinline__ double lane_read(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ float lane_read(float v, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, v)));
}

// N0: lane column (0 = xA, 1 = xA + 1) of the non-green site in rows of even
// offset from Y0; RED0: the non-green sites of those rows are red.  Rows of
// odd offset have the other column and the other colour.
//
// E: the strip's support reaches the edge of the active area.  The reference
// reads outside the area by whole-sample reflection, except the classifier,
// which reads gradients outside it as zero (MenonDemosaic.cpp:126-147).  The
// edge form loads reflected samples, so every symmetric stencil evaluated at
// a position outside the area (directional greens, ch / cv, red / blue)
// yields the value at its mirror image, which is what the reference reads
// there; the gradients of positions outside the area are set to zero; the
// selected green of a column outside the area is taken from its mirror
// column's lane, and the windows' rows outside the area from their mirror
// rows where the next stage reads them.  (Needs an area of at least 16 x 16,
// so that one reflection reaches every position a result depends on.)
template <int N0, int RED0, int OPS, bool E, bool REC, typename T>
struct Strip {
    static constexpr int ncol(int q) { return N0 ^ q; }
    static constexpr bool red_row(int q) { return (RED0 ^ q) != 0; }

    const StreamArgs& a;
    const int lane, xA;
    const int xa, xb;                 // load columns (clamped, or reflected, into the area)
    const int row_lo, row_hi;         // first and last area rows
    uint32_t e_all = 0;
    // Edge form: the lane's non-green column by row parity lies outside the
    // area, and the lane holding its mirror column.
    bool out_n[2] = {false, false};
    int mirror_lane[2] = {0, 0};

    // Windows, index k = row s + k (red / blue: row s + k - 1).
    float rn[7] = {}, rg[7] = {};     // normalised samples at the non-green / green site
    T gh[3] = {}, gv[3] = {};    // directional greens, rows s+2..s+4 (index k - 2)
    T dh[5] = {}, dhp[5] = {};   // horizontal gradients (own, lane L-1), rows s..s+4
    T dv[3] = {}, dvp[3] = {}, dvn[3] = {};   // vertical gradients (own, L-1, L+1), rows s..s+2
    T gr[3] = {};                // selected green at the non-green site, rows s..s+2
    bool hz[3] = {};                  // classifier decision (horizontal), rows s..s+2
    T red[3] = {}, blue[3] = {}; // red / blue at the green site, rows s-1..s+1
    // Codes at xA / xA + 1 of rows s+6 .. s+5+kPrefetch, each straight from
    // its 16-bit load (combining the two halves would make the loop wait for
    // the loads it has just issued).
    uint32_t ring_a[kPrefetch] = {}, ring_b[kPrefetch] = {};
    float pend[2][3] = {};            // op-list form: row s - 1's pixels, waiting for its pair
    bool pend_store = false;          // ... and whether row s - 1 is stored

    // FP32 output of the band's rows [yb, ye) through a buffer resource:
    // a store at an offset beyond the resource's size is discarded by the
    // hardware, so halo lanes and rows past the band store at kDrop and no
    // branch skips a store (a skipped store on one path makes every later
    // load wait for all earlier stores on the others).
    static constexpr uint32_t kDrop = 0x80000000u;
    int yb = 0, ye = 0;
    __amdgpu_buffer_rsrc_t rgb;

    __device__ Strip(const StreamArgs& a_, int lane_, int xA_, int yb_, int ye_)
        : a(a_), lane(lane_), xA(xA_),
          xa(E ? reflect_in(xA_, a_.ax, a_.ax + a_.aw - 1) : min(max(xA_, a_.ax), a_.ax + a_.aw - 1)),
          xb(E ? reflect_in(xA_ + 1, a_.ax, a_.ax + a_.aw - 1) : min(max(xA_ + 1, a_.ax), a_.ax + a_.aw - 1)),
          row_lo(a_.ay), row_hi(a_.ay + a_.ah - 1), yb(yb_), ye(ye_) {
        float* base = a_.out ? a_.out + (size_t)(yb_ - a_.vy) * a_.ostride * 3 : nullptr;
        const int bytes = a_.out ? (int)((size_t)(ye_ - yb_) * a_.ostride * 12) : 0;
        rgb = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, bytes, 0x00020000);
        if constexpr (E) {
            const int x0 = xA_ - 2 * lane_;   // lane 0's first column
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int x = xA_ + ncol(q);
                out_n[q] = x < a_.ax || x >= a_.ax + a_.aw;
                const int m = out_n[q] ? reflect_in(x, a_.ax, a_.ax + a_.aw - 1) : x;
                mirror_lane[q] = min(max((m - x0) >> 1, 0), 63);
            }
        }
    }

    __device__ __forceinline__ bool row_out(int y) const { return y < row_lo || y > row_hi; }

    // Byte offsets of the two pixels of row s in the band's output, kDrop
    // where the pixel is not one of the rectangle's.
    __device__ __forceinline__ void offsets(int s, bool row_ok, uint32_t& oa, uint32_t& ob) const {
        const bool ok = row_ok && lane >= 4 && lane < 60;
        const uint32_t o = (uint32_t)(((s - yb) * (int)a.ostride + (xA - a.vx)) * 12);
        oa = ok && xA < a.X1 ? o : kDrop;
        ob = ok && xA + 1 < a.X1 ? o + 12u : kDrop;
    }

    __device__ __forceinline__ void store3(const float* v, uint32_t off) const {
        typedef __attribute__((ext_vector_type(3))) unsigned u3;
        const u3 d = {__builtin_bit_cast(unsigned, v[0]), __builtin_bit_cast(unsigned, v[1]),
                      __builtin_bit_cast(unsigned, v[2])};
        __builtin_amdgcn_raw_buffer_store_b96(d, rgb, off, 0, 0);
    }

    __device__ __forceinline__ void load_row(int y, uint32_t& ca, uint32_t& cb) const {
        y = E ? reflect_in(y, row_lo, row_hi) : min(max(y, row_lo), row_hi);
        const uint16_t* p = a.samples + (size_t)y * a.stride;
        ca = p[xa];
        cb = p[xb];
    }

    // Normalised sample (RawEngine.cpp:89-97) of a code in a row of offset
    // parity q, lane column c.
    __device__ __forceinline__ float norm(uint32_t code, int q, int c) const {
        return normalize_exact((float)code, a.blk[q][c], a.rcp[q][c], a.den[q][c]);
    }

    // Stages of a row step; a band's first warm-up steps skip those whose
    // rows no later stage of the band reads.
    enum : int { ST_DV = 1, ST_CLS = 2, ST_CHR = 4, ST_ALL = 7 };

    // One row step; Q = offset parity of the final row s.  `final`: form
    // row s's pixels (the rows of a band's row groups); `store`: row s is a
    // row of the rectangle.
    template <int Q, int ST = ST_ALL>
    __device__ __forceinline__ void step(int s, bool final, bool store) {
        // ---- raw row s + 6 (offset parity Q) ------------------------------
        {
            const uint32_t ca = ring_a[0], cb = ring_b[0];
#pragma unroll
            for (int k = 0; k + 1 < kPrefetch; ++k) { ring_a[k] = ring_a[k + 1]; ring_b[k] = ring_b[k + 1]; }
            load_row(s + 6 + kPrefetch, ring_a[kPrefetch - 1], ring_b[kPrefetch - 1]);
            const float fa = norm(ca, Q, 0), fb = norm(cb, Q, 1);
            rn[6] = ncol(Q) ? fb : fa;
            rg[6] = ncol(Q) ? fa : fb;
        }
        // ---- directional greens, ch, dh at row s + 4 (MenonDemosaic.cpp:107-119) ----
        T gh4, gv4, dh4, dhp4;
        {
            const T own = rn[4];
            const float np = prev_f(rn[4]), nn = next_f(rn[4]);
            float gl, gright;
            if (ncol(Q) == 0) { gl = prev_f(rg[4]); gright = rg[4]; }
            else { gl = rg[4]; gright = next_f(rg[4]); }
            gh4 = T(0.5) * ((T)gl + (T)gright) + T(0.5) * own - T(0.25) * ((T)np + (T)nn);
            gv4 = T(0.5) * ((T)rg[3] + (T)rg[5]) + T(0.5) * own - T(0.25) * ((T)rn[2] + (T)rn[6]);
            const T ch = own - gh4;
            dh4 = abs_t(ch - next_t(ch));          // dh = |ch(x) - ch(x + 2)|
            if constexpr (E) {
                if (out_n[Q] || row_out(s + 4)) dh4 = 0.0;   // read as zero outside the area
            }
            dhp4 = prev_t(dh4);
        }
        gh[2] = gh4;
        gv[2] = gv4;
        dh[4] = dh4;
        dhp[4] = dhp4;
        // ---- cv, dv at row s + 2: dv = |cv(y) - cv(y + 2)| ------------------
        if constexpr ((ST & ST_DV) != 0) {
            const T cv2 = (T)rn[2] - gv[0];
            const T cv4 = (T)rn[4] - gv[2];
            T d = abs_t(cv2 - cv4);
            if constexpr (E) {
                if (out_n[Q] || row_out(s + 2)) d = 0.0;
            }
            dv[2] = d;
            dvp[2] = prev_t(d);
            dvn[2] = next_t(d);
        }
        // ---- classifier at row s + 2 (MenonDemosaic.cpp:126-147) ---------
        // Terms in the reference's order, weights {1,1,1,3,3,1,1,1}; the
        // gradients are >= +0, so the first term equals 0 + term.
        if constexpr ((ST & ST_CLS) != 0) {
            constexpr int n = ncol(Q);
            T h = dh[4];                      // ( 0,  2)
            h += dhp[4];                           // (-2,  2)
            h += n == 0 ? dhp[3] : dh[3];          // (-1,  1)
            h += T(3.0) * dh[2];                      // ( 0,  0)
            h += T(3.0) * dhp[2];                     // (-2,  0)
            h += n == 0 ? dhp[1] : dh[1];          // (-1, -1)
            h += dh[0];                            // ( 0, -2)
            h += dhp[0];                           // (-2, -2)
            T v = dvn[2];                     // ( 2,  0)
            v += dvn[0];                           // ( 2, -2)
            v += n == 0 ? dv[1] : dvn[1];          // ( 1, -1)
            v += T(3.0) * dv[2];                      // ( 0,  0)
            v += T(3.0) * dv[0];                      // ( 0, -2)
            v += n == 0 ? dvp[1] : dv[1];          // (-1, -1)
            v += dvp[2];                           // (-2,  0)
            v += dvp[0];                           // (-2, -2)
            const bool horizontal = v >= h;
            hz[2] = horizontal;
            gr[2] = horizontal ? gh[0] : gv[0];
            if constexpr (REC) record<Q>(s + 2, horizontal);
            if constexpr (E) gr[2] = lane_read(gr[2], out_n[Q] ? mirror_lane[Q] : lane);
        }
        // ---- red / blue at the green site of row s + 1 (MenonDemosaic.cpp:148-161) ----
        if constexpr ((ST & ST_CHR) != 0) {
            constexpr int q1 = Q ^ 1;
            if constexpr (E) {
                // The green of a row outside the area is its mirror row's.
                if (s + 1 == row_lo) gr[0] = gr[2];
                if (s + 1 == row_hi) gr[2] = gr[0];
            }
            const T own = rg[1];
            float rl, rr;
            T gl, grt;
            if (ncol(q1) == 0) { rl = rn[1]; gl = gr[1]; rr = next_f(rn[1]); grt = next_t(gr[1]); }
            else { rl = prev_f(rn[1]); gl = prev_t(gr[1]); rr = rn[1]; grt = gr[1]; }
            const T ch = own + T(0.5) * ((T)rl + (T)rr) - T(0.5) * (gl + grt);
            const T cv = own + T(0.5) * ((T)rn[0] + (T)rn[2]) - T(0.5) * (gr[0] + gr[2]);
            red[2] = red_row(q1) ? ch : cv;
            blue[2] = red_row(q1) ? cv : ch;
        }
        // ---- final RGB of row s (MenonDemosaic.cpp:162-173) ----------------
        if (final) finish<Q>(s, store);
        // ---- slide the windows ----------------------------------------------
#pragma unroll
        for (int k = 0; k < 6; ++k) { rn[k] = rn[k + 1]; rg[k] = rg[k + 1]; }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            gh[k] = gh[k + 1]; gv[k] = gv[k + 1];
            dv[k] = dv[k + 1]; dvp[k] = dvp[k + 1]; dvn[k] = dvn[k + 1];
            gr[k] = gr[k + 1]; hz[k] = hz[k + 1];
            red[k] = red[k + 1]; blue[k] = blue[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) { dh[k] = dh[k + 1]; dhp[k] = dhp[k + 1]; }
    }

    // The decision of row c (a band row) into the record: the strip's valid
    // lanes hold 56 consecutive non-green sites of the row, which touch at
    // most three 32-site words; lanes 0..2 OR in one word each (a word's
    // other sites belong to the neighbouring strip).
    template <int Q>
    __device__ __forceinline__ void record(int c, bool horizontal) {
        if (c < yb || c >= ye) return;
        const bool valid = lane >= 4 && lane < 60 && xA + ncol(Q) < a.X1;
        const uint64_t m = __ballot(horizontal && valid);
        const int k0 = (xA - 2 * lane - a.vx) >> 1;        // site index of lane 0
        const int first = (k0 + 4) >> 5, last = (k0 + 59) >> 5;
        const int word = first + lane;
        if (lane < 3 && word <= last) {
            const int shift = word * 32 - k0;
            const uint32_t bits = (uint32_t)(shift >= 0 ? m >> shift : m << -shift);
            if (bits) {
                const int Y = c - a.vy;
                atomicOr(a.decisions + ((size_t)(Y >> 5) * a.dtiles + (size_t)word) * 32 + (Y & 31), bits);
            }
        }
    }

    template <int Q>
    __device__ __forceinline__ void finish(int s, bool store) {
        constexpr int n = ncol(Q);
        if constexpr (E) {
            // Red / blue of a row outside the area are its mirror row's.
            if (s == row_lo) { red[0] = red[2]; blue[0] = blue[2]; }
            if (s == row_hi) { red[2] = red[0]; blue[2] = blue[0]; }
        }
        // Red / blue along the decision at the non-green site; the two
        // horizontal neighbours are green sites of row s, the two vertical
        // ones green sites of rows s -+ 1 in this column.
        T hl_r, hl_b, hr_r, hr_b;
        if (n == 0) { hl_r = prev_t(red[1]); hl_b = prev_t(blue[1]); hr_r = red[1]; hr_b = blue[1]; }
        else { hl_r = red[1]; hl_b = blue[1]; hr_r = next_t(red[1]); hr_b = next_t(blue[1]); }
        const bool horizontal = hz[0];
        const T mr = T(0.5) * (horizontal ? hl_r + hr_r : red[0] + red[2]);
        const T mb = T(0.5) * (horizontal ? hl_b + hr_b : blue[0] + blue[2]);
        const T own = rn[0];
        float px[2][3];
        float* P = px[n];
        float* G = px[1 - n];
        if (red_row(Q)) {
            P[0] = rn[0];
            P[2] = (float)(own + mb - mr);     // bb = rr + mean(blue) - mean(red)
        } else {
            P[2] = rn[0];
            P[0] = (float)(own + mr - mb);     // rr = bb + mean(red) - mean(blue)
        }
        P[1] = (float)gr[0];
        G[0] = (float)red[1];
        G[1] = rg[0];
        G[2] = (float)blue[1];

        if constexpr ((OPS & STREAM_OPS_LINEAR) != 0) {
            // WhiteBalanceNode, ExposureNode, CameraToWorkingNode
            // (RawEngine.cpp:408-548): each product and sum rounds to float.
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float r = px[p][0], g = px[p][1], b = px[p][2];
                if constexpr ((OPS & STREAM_LIN_GAINS) != 0) {
                    r *= a.gains[0]; g *= a.gains[1]; b *= a.gains[2];
                    r *= a.mult; g *= a.mult; b *= a.mult;
                }
                if constexpr ((OPS & STREAM_LIN_MATRIX) != 0) {
                    const float R = a.m[0] * r + a.m[1] * g + a.m[2] * b;
                    const float Gc = a.m[3] * r + a.m[4] * g + a.m[5] * b;
                    const float B = a.m[6] * r + a.m[7] * g + a.m[8] * b;
                    r = R; g = Gc; b = B;
                }
                px[p][0] = r; px[p][1] = g; px[p][2] = b;
            }
            uint32_t oa, ob;
            offsets(s, store, oa, ob);
            if (a.out) { store3(px[0], oa); store3(px[1], ob); }
        } else {
            // The device op lists, once per row pair: row s - 1 waits in
            // `pend` for row s.
            if constexpr (Q == 0) {
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int c = 0; c < 3; ++c) pend[p][c] = px[p][c];
                pend_store = store;
            } else {
                float q[4][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    q[0][c] = pend[0][c]; q[1][c] = pend[1][c]; q[2][c] = px[0][c]; q[3][c] = px[1][c];
                }
                uint32_t err[4] = {0u, 0u, 0u, 0u};
                apply_ops_px(*a.native, 0, as_const(*a.native).n, q, err);
                const OpList& R = *a.reduced;
                const uint32_t nr = as_const(R).n;
                const uint32_t tap = a.tap < nr ? a.tap : nr;
                apply_ops_px(R, 0, tap, q, err);
                const bool ok = lane >= 4 && lane < 60;
                const bool okA = ok && xA < a.X1, okB = ok && xA + 1 < a.X1;
                if (a.out) {
                    uint32_t oa, ob;
                    offsets(s - 1, pend_store, oa, ob);
                    store3(q[0], oa); store3(q[1], ob);
                    offsets(s, store, oa, ob);
                    store3(q[2], oa); store3(q[3], ob);
                }
                if (a.out8) {
                    apply_ops_px(R, tap, nr, q, err);
                    if (pend_store) store_u8(s - 1, q[0], q[1], okA, okB);
                    if (store) store_u8(s, q[2], q[3], okA, okB);
                }
                if (pend_store) e_all |= (okA ? err[0] : 0u) | (okB ? err[1] : 0u);
                if (store) e_all |= (okA ? err[2] : 0u) | (okB ? err[3] : 0u);
            }
        }
    }

    __device__ __forceinline__ void store_u8(int s, const float* pa, const float* pb, bool okA, bool okB) const {
        uint8_t* d = a.out8 + ((size_t)(s - a.vy) * a.ostride + (size_t)(xA - a.vx)) * 3;
        if (okA) { d[0] = quantize_u8(pa[0]); d[1] = quantize_u8(pa[1]); d[2] = quantize_u8(pa[2]); }
        if (okB) { d[3] = quantize_u8(pb[0]); d[4] = quantize_u8(pb[1]); d[5] = quantize_u8(pb[2]); }
    }
};

template <int N0, int RED0, int OPS, bool E, bool REC, typename T>
__device__ __forceinline__ void run_strip(const StreamArgs& a, int lane, int xA, int yb, int ye) {
    const int ye4 = yb + ((ye - yb + 3) & ~3);                // whole groups of four rows
    Strip<N0, RED0, OPS, E, REC, T> t(a, lane, xA, yb, ye);
    // Rows yb-6 .. yb-3 of the window before the first step (s = yb - 8
    // reads rows s+2 .. s+6), then the rows in flight.
#pragma unroll
    for (int k = 2; k < 6; ++k) {
        uint32_t ca, cb;
        t.load_row(yb - 8 + k, ca, cb);
        const int q = k & 1;
        const float fa = t.norm(ca, q, 0), fb = t.norm(cb, q, 1);
        t.rn[k] = (N0 ^ q) ? fb : fa;
        t.rg[k] = (N0 ^ q) ? fa : fb;
    }
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) t.load_row(yb - 2 + k, t.ring_a[k], t.ring_b[k]);
    // Eight warm-up rows (their final rows lie above the band; each runs
    // only the stages whose rows the band reads: dv from row yb - 4, the
    // classifier from yb - 2, red / blue from yb - 1), then the band, four
    // rows per iteration: the windows' registers rename within an iteration
    // and only the few values older than four rows move.
    using S = Strip<N0, RED0, OPS, E, REC, T>;
    t.template step<0, 0>(yb - 8, false, false);
    t.template step<1, 0>(yb - 7, false, false);
    t.template step<0, S::ST_DV>(yb - 6, false, false);
    t.template step<1, S::ST_DV>(yb - 5, false, false);
    t.template step<0, S::ST_DV | S::ST_CLS>(yb - 4, false, false);
    t.template step<1, S::ST_DV | S::ST_CLS>(yb - 3, false, false);
    t.template step<0, S::ST_ALL>(yb - 2, false, false);
    t.template step<1, S::ST_ALL>(yb - 1, false, false);
    for (int s = yb; s < ye4; s += 4) {
        t.template step<0>(s, true, true);
        t.template step<1>(s + 1, true, s + 1 < ye);
        t.template step<0>(s + 2, true, s + 2 < ye);
        t.template step<1>(s + 3, true, s + 3 < ye);
    }
    if (t.e_all && a.flags) atomicOr(a.flags, t.e_all);
}

template <int N0, int RED0, int OPS, bool REC = false, typename T = double>
__global__ void __launch_bounds__(kStreamThreads) menon_stream_kernel(const StreamArgs a) {
    const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kStreamThreads / 64) + (threadIdx.x >> 6)));
    if (wave >= a.wave_cols * a.bands) return;   // whole wavefronts leave together
    const int band = wave / a.wave_cols, wcol = wave - band * a.wave_cols;
    const int lane = (int)(threadIdx.x & 63);
    const int x0 = a.X0 + 112 * wcol - 8;                     // lane 0's first column
    const int xA = x0 + 2 * lane;
    const int yb = a.Y0 + band * a.band_rows;                 // even offset from Y0
    const int ye = min(yb + a.band_rows, a.Y1);
    // A strip whose columns or rows (with the 6-sample support of the band's
    // rows and the warm-up rows above them) come within reach of the area's
    // edge takes the edge form.
    const bool edge = x0 < a.ax + 8 || x0 + 128 > a.ax + a.aw - 8 || yb - 16 < a.ay || ye + 16 > a.ay + a.ah;
    if (edge) run_strip<N0, RED0, OPS, true, REC, T>(a, lane, xA, yb, ye);
    else run_strip<N0, RED0, OPS, false, REC, T>(a, lane, xA, yb, ye);
}

__global__ void zero_words(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

hipError_t launch_menon_stream(const StreamArgs& args, int n0, int red0, int ops, hipStream_t stream) {
    StreamArgs a = args;
    if (a.X1 <= a.X0 || a.Y1 <= a.Y0) return hipSuccess;
    a.wave_cols = (a.X1 - a.X0 + 111) / 112;
    const int H = a.Y1 - a.Y0;
    // 16-row bands: measured on the 24 MP frame (stream kernel alone, bands
    // of 8 / 12 / 16 / 20 / 24 / 32 / 48 / 72 rows: 122.5 / 110.9 / 106.5 /
    // 113.3 / 116.5 / 123.5 / 130.7 / 128.5 us).  Longer bands repeat fewer
    // warm-up rows but keep every wavefront resident for the whole launch in
    // the same phase of its row step, and issue fewer instructions per cycle.
    if (a.band_rows <= 0) a.band_rows = 16;
    a.bands = (H + a.band_rows - 1) / a.band_rows;
    const int waves = a.wave_cols * a.bands;
    const unsigned blocks = (unsigned)((waves + kStreamThreads / 64 - 1) / (kStreamThreads / 64));
    const dim3 grid(blocks), block(kStreamThreads);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, 0, stream, a); };
    const bool fp32 = a.fp32 && ops != STREAM_OPS_LISTS;
    if (a.decisions) {
        // The record's words are OR-ed into: clear them first.
        const size_t words = (size_t)a.dtiles * (size_t)((a.Y1 - a.Y0 + 31) / 32) * 32;
        hipLaunchKernelGGL(zero_words, dim3((unsigned)std::min<size_t>((words + 255) / 256, 4096)), dim3(256), 0, stream,
                           a.decisions, words);
        if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
#define RAWOPS_STREAM_REC(N, R, T)                                                             \
    switch (ops) {                                                                             \
    case STREAM_OPS_LINEAR: go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR, true, T>); break;  \
    case STREAM_OPS_LINEAR | STREAM_LIN_GAINS:                                                 \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS, true, T>); break;   \
    case STREAM_OPS_LINEAR | STREAM_LIN_MATRIX:                                                \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_MATRIX, true, T>); break;  \
    default:                                                                                   \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS | STREAM_LIN_MATRIX, true, T>); \
        break;                                                                                 \
    }
#define RAWOPS_STREAM_REC_T(T)                                                                 \
        if (n0) {                                                                              \
            if (red0) { RAWOPS_STREAM_REC(1, 1, T) } else { RAWOPS_STREAM_REC(1, 0, T) }       \
        } else {                                                                               \
            if (red0) { RAWOPS_STREAM_REC(0, 1, T) } else { RAWOPS_STREAM_REC(0, 0, T) }       \
        }
        if (fp32) { RAWOPS_STREAM_REC_T(float) } else { RAWOPS_STREAM_REC_T(double) }
#undef RAWOPS_STREAM_REC_T
#undef RAWOPS_STREAM_REC
        return hipGetLastError();
    }
    if (fp32) {
#define RAWOPS_STREAM_F32(N, R)                                                                \
    switch (ops) {                                                                             \
    case STREAM_OPS_LINEAR: go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR, false, float>); break; \
    case STREAM_OPS_LINEAR | STREAM_LIN_GAINS:                                                 \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS, false, float>); break; \
    case STREAM_OPS_LINEAR | STREAM_LIN_MATRIX:                                                \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_MATRIX, false, float>); break; \
    default:                                                                                   \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS | STREAM_LIN_MATRIX, false, float>); \
        break;                                                                                 \
    }
        if (n0) {
            if (red0) { RAWOPS_STREAM_F32(1, 1) } else { RAWOPS_STREAM_F32(1, 0) }
        } else {
            if (red0) { RAWOPS_STREAM_F32(0, 1) } else { RAWOPS_STREAM_F32(0, 0) }
        }
#undef RAWOPS_STREAM_F32
        return hipGetLastError();
    }
#define RAWOPS_STREAM_FORMS(N, R)                                                              \
    switch (ops) {                                                                             \
    case STREAM_OPS_LINEAR: go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR>); break;           \
    case STREAM_OPS_LINEAR | STREAM_LIN_GAINS:                                                 \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS>); break;            \
    case STREAM_OPS_LINEAR | STREAM_LIN_MATRIX:                                                \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_MATRIX>); break;           \
    case STREAM_OPS_LINEAR | STREAM_LIN_GAINS | STREAM_LIN_MATRIX:                             \
        go(menon_stream_kernel<N, R, STREAM_OPS_LINEAR | STREAM_LIN_GAINS | STREAM_LIN_MATRIX>); \
        break;                                                                                 \
    default: go(menon_stream_kernel<N, R, STREAM_OPS_LISTS>); break;                           \
    }
    if (n0) {
        if (red0) { RAWOPS_STREAM_FORMS(1, 1) } else { RAWOPS_STREAM_FORMS(1, 0) }
    } else {
        if (red0) { RAWOPS_STREAM_FORMS(0, 1) } else { RAWOPS_STREAM_FORMS(0, 0) }
    }
#undef RAWOPS_STREAM_FORMS
    return hipGetLastError();
}

}  // namespace rawops
