// Adjoint (reverse-mode) kernels of the render chain
//     out = M * (m * (g (.) demosaic(raw)))
// (WhiteBalanceNode, ExposureNode, CameraToWorkingNode applied after the
// demosaic, RawEngine.cpp:408-548) for gfx950.
//
// Given dL/dout over a viewport, one launch produces dL/d(normalised raw) over
// the whole active area and accumulates dL/d(gains, multiplier, matrix).
//   * Bilinear: the transpose of the same-colour mean, i.e. a remosaic
//     scatter, restricted per tile so each workgroup owns its raw outputs.
//   * Menon: the transpose of the branch actually taken.  With the classifier
//     decisions frozen (recomputed here exactly as the forward makes them),
//     the demosaic is linear in the raw samples; its adjoint walks the stages
//     backwards: final chroma -> green-site chroma -> directional green.
//     Each stage's transpose is taken as a gather (a fixed neighbour stencil
//     over the recorded decisions); reads that the forward mirrors at the
//     active-area edge are gathered from the mirrored positions.
// A Menon workgroup recomputes the forward over its tile grown by 4 (the
// adjoint's reach) and writes only its own tile, so neither LDS nor global
// atomics touch the raw gradient.
#include "device_common.hpp"
#include "menon_fwd.hpp"
#include "raw_tile.hpp"
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kNT = 512;
// Menon adjoint tile height (64-wide tiles; 24 MP frame, recorded /
// recomputing adjoint, ms): 20 rows 0.279 / 0.331, 24: 0.256 / 0.386,
// 28: 0.243 / 0.347, 32: 0.228 / 0.343, 40: 0.258 / 0.507, 48: 0.237 /
// 0.462.  One height for both kernels, so that
// a site is gathered by the same (interior or border) tile code in both and
// their raw gradients agree bit for bit.
constexpr int kAdjTH = 32;

struct BackwardArgs {
    SensorArgs s;
    int vx, vy, vw, vh;        // viewport (native sensor coordinates)
    float g[3], m, M[9];
    float K[9];                // K[c][r] = g_c m M[r][c] (chain back to the demosaic output)
    const float* gout;         // viewport RGB
    uint32_t gstride;          // pixels per gout row
    float* graw;               // active-area plane, aw x ah
    double* gparams;           // 13: gains[3], multiplier, matrix[9] (row-major)
    int tiles_x, tiles_y;      // tiles of the launch
    int tile0_x, tile0_y;      // first tile (in tiles from the area origin)
    int r0, r1, c0, c1;        // launch-relative tile rows / columns of the interior block
    const uint32_t* decisions; // recorded Menon decisions (bits, relative to the recorded viewport) or null
    uint32_t dtiles;           // 64 x 32 tiles per recorded row (32 decision words each)
    int rx, ry, rw, rh;        // the recorded viewport
    int ir0, ir1, ic0, ic1;    // within the recorded block: its tiles that need no mirroring
};

// dL/dout at a sensor pixel (zero outside the viewport).
__device__ __forceinline__ bool load_gout(const BackwardArgs& a, int sx, int sy, float go[3]) {
    const int lx = sx - a.vx, ly = sy - a.vy;
    if (lx < 0 || ly < 0 || lx >= a.vw || ly >= a.vh) return false;
    const float* p = a.gout + ((size_t)ly * a.gstride + (size_t)lx) * 3;
    go[0] = p[0]; go[1] = p[1]; go[2] = p[2];
    return true;
}

// gd_c = g_c m (M^T gout)_c: gradient at the demosaic output, through the
// combined matrix K[c][r] = g_c m M[r][c] (formed in FP64 on the host,
// rounded once).
__device__ __forceinline__ void chain_back(const BackwardArgs& a, const float go[3], float gd[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gd[c] = a.K[3 * c] * go[0] + a.K[3 * c + 1] * go[1] + a.K[3 * c + 2] * go[2];
}

// Parameter gradients.  With u = M^T gout and out = M (m g (.) d), every
// parameter gradient is a fixed combination of the nine sums
//     P[r][c] = sum over pixels of gout_r d_c :
//     dL/dg_c    = m sum_r M[r][c] P[r][c]
//     dL/dm      = sum_c g_c sum_r M[r][c] P[r][c]
//     dL/dM[r][c] = m g_c P[r][c]
// so a pixel adds nine products and the combinations are taken once per
// workgroup.  A thread's few products (at most eight pixels) are summed by
// FP32 fused multiply-adds (relative error ~1e-7 of the partial sum); the
// partials are summed in FP64 from there on.
__device__ __forceinline__ void param_grads(const float d[3], const float go[3], float (&acc)[9]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r * 3 + c] = __builtin_fmaf(go[r], d[c], acc[r * 3 + c]);
}

// x plus the value DPP control CTRL moves into this lane (rows outside
// ROW_MASK, and lanes without a source, add zero).  A control that gives
// every lane a source (ROW_MASK 0xf) needs no zeroed old value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    uint32_t lo, hi;
    if constexpr (ROW_MASK == 0xf) {
        lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, false);
        hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    } else {
        lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
        hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
    }
    return x + __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the wave's 64 lanes, complete in lane 63: pairs and quads
// (quad permutes), eight and sixteen lanes (row half mirror, row mirror),
// then rows 1 and 3 take row 0 and 2 (row broadcast 15) and rows 2 and 3
// take rows 0 + 1 (row broadcast 31).  Register moves only.
__device__ __forceinline__ double wave_sum63(double x) {
    x = dpp_add<0xb1, 0xf>(x);    // quad_perm [1, 0, 3, 2]
    x = dpp_add<0x4e, 0xf>(x);    // quad_perm [2, 3, 0, 1]
    x = dpp_add<0x141, 0xf>(x);   // row_half_mirror
    x = dpp_add<0x140, 0xf>(x);   // row_mirror
    x = dpp_add<0x142, 0xa>(x);   // row_bcast:15
    x = dpp_add<0x143, 0xc>(x);   // row_bcast:31
    return x;
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The workgroup's nine sums P[r][c] = sum gout_r d_c -> the 13 parameter
// gradients (added to a.gparams):
//   k < 3:   m sum_r M[r][k] P[r][k]             (dL/d gain k)
//   k == 3:  sum_c g_c sum_r M[r][c] P[r][c]     (dL/d exposure multiplier)
//   k >= 4:  m g_c P[k - 4], c = (k - 4) mod 3   (dL/d M)
// Per wave by DPP, then across waves in wave 0 (its lanes 0..8 sum one P
// each over the waves, in wave order; the P are then broadcast).
__device__ __forceinline__ void flush_params(const BackwardArgs& a, const float (&part)[9], double* red /* [waves][9] LDS */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = wave_sum63((double)part[i]);
    if (lane == 63)
        for (int i = 0; i < 9; ++i) red[wave * 9 + i] = acc[i];
    __syncthreads();
    if (wave != 0) return;
    double t = 0;
    if (lane < 9) {
#pragma unroll
        for (int w = 0; w < kNT / 64; ++w) t += red[w * 9 + lane];
    }
    double P[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) P[i] = readlane_f64(t, i);
    const double m = (double)a.m;
    double S[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        S[c] = (double)a.M[c] * P[c] + (double)a.M[3 + c] * P[3 + c] + (double)a.M[6 + c] * P[6 + c];
    if (lane >= 13) return;
    const int k = lane;
    double v;
    if (k < 3) {
        v = m * (k == 0 ? S[0] : k == 1 ? S[1] : S[2]);
    } else if (k == 3) {
        v = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) v += (double)a.g[c] * S[c];
    } else {
        const int c = (k - 4) % 3;
        const double g = c == 0 ? (double)a.g[0] : c == 1 ? (double)a.g[1] : (double)a.g[2];
        double pk = P[0];
#pragma unroll
        for (int i = 1; i < 9; ++i) pk = (k - 4 == i) ? P[i] : pk;
        v = m * g * pk;
    }
    atomicAdd(a.gparams + k, v);
}

// Border-first tile order: the tiles outside the interior block [r0, r1) x
// [c0, c1) (slower: mirrored reads, scatter adjoint) take the first block
// indices so they start with the launch instead of trailing it.
// Block b of a TY x TX tile grid, the tiles outside the block [r0, r1) x
// [c0, c1) first (a bijection; an empty block lists the grid row-major).
__device__ __forceinline__ void border_first(int b, int TX, int TY, int r0, int r1, int c0, int c1, int& r, int& c) {
    const int top = r0 * TX, bottom = (TY - r1) * TX, side = c0 + (TX - c1), middle = (r1 - r0) * side;
    if (b < top) {
        r = b / TX; c = b - r * TX;
    } else if ((b -= top) < bottom) {
        r = r1 + b / TX; c = b - (b / TX) * TX;
    } else if ((b -= bottom) < middle) {
        r = r0 + b / side;
        const int j = b - (b / side) * side;
        c = j < c0 ? j : c1 + (j - c0);
    } else {
        b -= middle;
        const int ci = c1 - c0;
        r = r0 + b / ci; c = c0 + (b - (b / ci) * ci);
    }
}

__device__ __forceinline__ void tile_of_border_first(const BackwardArgs& a, int& tx, int& ty, int TW, int TH) {
    int r, c;
    border_first((int)blockIdx.x, a.tiles_x, a.tiles_y, a.r0, a.r1, a.c0, a.c1, r, c);
    tx = a.s.ax + (a.tile0_x + c) * TW;
    ty = a.s.ay + (a.tile0_y + r) * TH;
}

__device__ __forceinline__ void tile_of(const BackwardArgs& a, int& tx, int& ty, int TW, int TH) {
    tx = a.s.ax + (a.tile0_x + (int)(blockIdx.x % a.tiles_x)) * TW;
    ty = a.s.ay + (a.tile0_y + (int)(blockIdx.x / a.tiles_x)) * TH;
}

// --------------------------------------------------------------- Menon --
//
// The Menon adjoint is the transpose of the branch the classifier takes: with
// the decisions frozen, the demosaic is linear in the raw samples.  A
// workgroup owns one tile of dL/draw and works on its extended tile (tile
// +-4, the adjoint's reach) in coordinates (u, v) = sensor - (ex, ey); the
// tile is [4, 4+TW) x [4, 4+TH).  Every stage's transpose is a gather -- the
// adjoint reaching a site is a fixed stencil over its neighbours and their
// decisions -- so no LDS or global atomics touch the raw gradient:
//   pass 1: gd = dL/d demosaic output at every site; non-green sites publish
//           the channel the final chroma step mixes, split by the site's
//           decision into POh (horizontal) and POv (vertical); parameter
//           gradients of the tile's pixels
//   pass 2: green sites: red/blue adjoints AR, AB = own gd -+ 0.5 of the
//           covering neighbours' PO (horizontal ones from POh, vertical ones
//           from POv)
//   pass 3: non-green sites: green adjoint AG = gd_g - 0.5 sum of the four
//           neighbours' AR (red sites) or AB (blue sites), split into
//           AGh / AGv by the site's decision
//   pass 4: tile sites: raw adjoint through the directional-green stencil.
// Adjoint planes are compacted per site class: index v (EW/2) + (u >> 1).
//
// The forward values the passes need -- the decisions over tile +-4 and the
// demosaic output of the tile's pixels -- come either from a recompute of
// the forward over the extended tile (menon_backward), or from decisions a
// recording forward wrote (menon_backward_recorded: only the selected green
// and the green-site chroma are recomputed, from the samples of tile +-4).

template <int N0>
__device__ constexpr int ng_x(int b) { return (b & 1) ? 1 - N0 : N0; }   // x parity of non-green sites, row parity b

// Upstream gradients of a thread's row-pair items, fetched early.  Item
// (qr, k) holds, in each row v = 2 qr + b, the non-green site 2k + ng_x(b)
// and the green site next to it; both have compacted index v (EW/2) + k.
template <int J>
struct Upstream {
    float goN[J][2][3], goG[J][2][3];
    uint32_t hasN, hasG;    // bit 2j + b
};

// Branch-free: every load is issued, from a position clamped into the
// viewport (and the last item), so that the loads stay in flight together;
// the has-bits say which are real.
template <int EW, int EH, int N0, int J>
__device__ __forceinline__ void fetch_upstream(const BackwardArgs& a, int ex, int ey, Upstream<J>& up) {
    constexpr int HWE = EW / 2, NI = HWE * (EH / 2);
    up.hasN = up.hasG = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        const int ic = min(it, NI - 1);
        const int qr = ic / HWE, k = ic - qr * HWE;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, un = 2 * k + ng_x<N0>(b);
            const int ly = ey + v - a.vy, lyc = min(max(ly, 0), a.vh - 1);
            const bool row = it < NI && ly == lyc;
            const float* prow = a.gout + (size_t)lyc * a.gstride * 3;
            auto fetch = [&](int u, float go[3]) -> bool {
                const int lx = ex + u - a.vx, lxc = min(max(lx, 0), a.vw - 1);
                const float* p = prow + (size_t)lxc * 3;
                // read once: non-temporal, so the stream does not evict the samples
                go[0] = __builtin_nontemporal_load(p); go[1] = __builtin_nontemporal_load(p + 1);
                go[2] = __builtin_nontemporal_load(p + 2);
                return row && lx == lxc;
            };
            if (fetch(un, up.goN[j][b])) up.hasN |= 1u << (2 * j + b);
            if (fetch(un ^ 1, up.goG[j][b])) up.hasG |= 1u << (2 * j + b);
        }
    }
}

// The four passes on an extended tile all of whose taps lie inside the
// active area, as row-pair items (every neighbour a compile-time offset, a
// thread's per-site values kept in registers across passes).  hz: decisions
// (index v (EW/2) + (u >> 1)); red_rows bit b: the non-green sites of rows of
// parity b are red; dn(u, v, b, d) / dg(u, v, b, d): demosaic output of the
// tile's non-green / green pixel (u, v) of row parity b.  POh/POv are written
// in pass 1, AR/AB in pass 2, AGh/AGv in pass 3; red is the parameter
// reduction buffer (written last).
template <int TW, int TH, int N0, int J, class DN, class DG>
__device__ __forceinline__ void interior_gather(const BackwardArgs& a, const SensorArgs& s, int tx, int ty,
                                                const Upstream<J>& up, const uint8_t* hz, uint32_t red_rows,
                                                float* POh, float* POv, float* AR, float* AB, float* AGh, float* AGv,
                                                double* red, DN dn, DG dg) {
    constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NI = HWE * (EH / 2);
    static_assert(J == (NI + kNT - 1) / kNT, "item count");
    const bool red0 = (red_rows & 1u) != 0, red1 = (red_rows & 2u) != 0;
    auto red_row = [&](int b) { return b ? red1 : red0; };
    auto tile_item = [](int qr, int k) { return qr >= 2 && qr < 2 + TH / 2 && k >= 2 && k < 2 + TW / 2; };
    // per item and row: green-channel gd and gd_r + gd_b of the non-green
    // site, gd of the green site, raw partials of both
    float pg[J][2], rn[J][2], gg[J][2][3], rpg[J][2], rpn[J][2];
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

    // pass 1
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
        const bool mine = tile_item(qr, k);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, un = 2 * k + ng_x<N0>(b), ug = un ^ 1, c = v * HWE + k;
            const bool rs = red_row(b);
            const bool hn = (up.hasN >> (2 * j + b)) & 1u, hg = (up.hasG >> (2 * j + b)) & 1u;
            float gd[3];
            if (hn) chain_back(a, up.goN[j][b], gd);
            else gd[0] = gd[1] = gd[2] = 0.f;
            const bool h = hz[c] != 0;
            const float po = rs ? gd[2] : gd[0];
            POh[c] = h ? po : 0.f;
            POv[c] = h ? 0.f : po;
            pg[j][b] = gd[1];
            rn[j][b] = gd[0] + gd[2];           // the own sample feeds both r and b
            if (hg) chain_back(a, up.goG[j][b], gg[j][b]);
            else gg[j][b][0] = gg[j][b][1] = gg[j][b][2] = 0.f;
            if (mine) {
                float d[3];
                if (hn) { dn(un, v, b, d); param_grads(d, up.goN[j][b], acc); }
                if (hg) { dg(ug, v, b, d); param_grads(d, up.goG[j][b], acc); }
            }
        }
    }
    __syncthreads();
    // pass 2: green sites with all four neighbours inside the extended tile
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, nb = ng_x<N0>(b), ug = 2 * k + 1 - nb, c = v * HWE + k;
            if (ug < 1 || ug >= EW - 1 || v < 1 || v >= EH - 1) continue;
            // R neighbour: B = r + mean(B) - mean(R); B neighbour: R = b + mean(R) - mean(B)
            const float sh = 0.5f * (POh[c - nb] + POh[c + 1 - nb]);      // row neighbours (row parity b)
            const float sv = 0.5f * (POv[c - HWE] + POv[c + HWE]);        // column neighbours (parity 1 - b)
            const float th = red_row(b) ? sh : -sh, tv = red_row(b ^ 1) ? sv : -sv;
            const float ar = gg[j][b][0] - th - tv;
            const float ab = gg[j][b][2] + th + tv;
            AR[c] = ar;
            AB[c] = ab;
            rpg[j][b] = ar + ab + gg[j][b][1];
        }
    }
    __syncthreads();
    // pass 3: non-green sites of tile +-2
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, nb = ng_x<N0>(b), un = 2 * k + nb, c = v * HWE + k;
            if (un < 2 || un >= EW - 2 || v < 2 || v >= EH - 2) continue;
            const float* A = red_row(b) ? AR : AB;
            const float S = (A[c - 1 + nb] + A[c + nb]) + (A[c - HWE] + A[c + HWE]);
            const float ag = pg[j][b] - 0.5f * S;
            const bool h = hz[c] != 0;
            AGh[c] = h ? ag : 0.f;
            AGv[c] = h ? 0.f : ag;
            rpn[j][b] = rn[j][b] + 0.5f * S;
        }
    }
    __syncthreads();
    // pass 4: raw adjoint of the tile through the directional-green stencil
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
        if (!tile_item(qr, k)) continue;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, nb = ng_x<N0>(b), un = 2 * k + nb, ug = un ^ 1, c = v * HWE + k;
            const float rgv = rpg[j][b] + 0.5f * (AGh[c - nb] + AGh[c + 1 - nb]) + 0.5f * (AGv[c - HWE] + AGv[c + HWE]);
            const float rnv = rpn[j][b] + 0.5f * (AGh[c] + AGv[c]) - 0.25f * (AGh[c - 1] + AGh[c + 1]) -
                              0.25f * (AGv[c - 2 * HWE] + AGv[c + 2 * HWE]);
            if (a.graw) {
                float* row = a.graw + (size_t)(ty + v - 4 - s.ay) * s.aw + (size_t)(tx - 4 - s.ax);
                __builtin_nontemporal_store(rnv, row + un);
                __builtin_nontemporal_store(rgv, row + ug);
            }
        }
    }
    flush_params(a, acc, red);
}

// The four passes on an extended tile that reaches the active-area edge.
// The forward mirrors reads there (whole-sample reflection m, per
// coordinate): a stage output at p reads its input at m(p + k) for each tap
// k, so the adjoint reaching a target q gathers over the virtual positions
// V(q) = {z : m(z) = q} within a tap's reach (2) of the area -- q itself and
// its reflections across nearby edges -- applying the interior stencil at
// each z and counting only sources inside the area.  Per site, with
// upstream gradients re-read where a pass needs them; RP holds the tile's raw
// partials (written in passes 2 and 3).  Arguments as interior_gather.
template <int TW, int TH, int N0, class DN, class DG>
__device__ __forceinline__ void border_gather(const BackwardArgs& a, const SensorArgs& s, int tx, int ty,
                                              const uint8_t* hz, uint32_t red_rows, float* POh, float* POv,
                                              float* AR, float* AB, float* AGh, float* AGv, float* RP, double* red,
                                              DN dn, DG dg) {
    constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NC = HWE * EH;
    const int ex = tx - 4, ey = ty - 4;
    auto ci = [](int u, int v) { return v * HWE + (u >> 1); };
    auto in_tile = [](int u, int v) { return u >= 4 && u < 4 + TW && v >= 4 && v < 4 + TH; };
    auto in_ext = [](int u, int v) { return u >= 0 && u < EW && v >= 0 && v < EH; };
    const int ax = s.ax, ay = s.ay, aw = s.aw, ah = s.ah;
    auto in_a = [&](int u, int v) {
        const int x = ex + u, y = ey + v;
        return x >= ax && y >= ay && x < ax + aw && y < ay + ah;
    };
    auto src = [&](int u, int v) { return in_ext(u, v) && in_a(u, v); };    // a computed stage output
    auto red_at = [&](int v) { return ((red_rows >> (v & 1)) & 1u) != 0; };  // non-green sites of row v
    auto gd_at = [&](int u, int v, float gd[3], float go[3]) -> bool {
        if (!load_gout(a, ex + u, ey + v, go)) { gd[0] = gd[1] = gd[2] = 0.f; return false; }
        chain_back(a, go, gd);
        return true;
    };
    // Per coordinate: the position itself and the outside positions within 2
    // of the area that reflect onto it (at most two, kept in scalars).
    auto axis = [](int q, int lo, int n, int& z1, int& z2) {
        int cnt = 1;
        z1 = z2 = q;
        auto take = [&](int c) {
            // whole-sample reflection as its periodic fold (period 2(n-1)), loop-free
            const int period = 2 * (n - 1);
            int t = period > 0 ? (c - lo) % period : 0;
            if (t < 0) t += period;
            const int m = lo + (t < n ? t : period - t);
            if (cnt < 3 && m == q) {
                if (cnt == 1) z1 = c;
                else z2 = c;
                ++cnt;
            }
        };
        take(lo - 2);
        take(lo - 1);
        take(lo + n);
        take(lo + n + 1);
        return cnt;
    };
    auto virtuals = [&](int u, int v, auto&& fn) {
        const int qx = ex + u, qy = ey + v;
        int x1, x2, y1, y2;
        const int nx = axis(qx, ax, aw, x1, x2), ny = axis(qy, ay, ah, y1, y2);
        for (int i = 0; i < ny; ++i) {
            const int zy = i == 0 ? qy : (i == 1 ? y1 : y2);
            for (int k = 0; k < nx; ++k) {
                const int zx = k == 0 ? qx : (k == 1 ? x1 : x2);
                fn(zx - ex, zy - ey);
            }
        }
    };
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

    // pass 1: every site of the extended tile inside the area, as row-pair
    // items (compile-time CFA parities)
    for (int it = threadIdx.x; it < NC / 2; it += kNT) {
        const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int gsite = 0; gsite < 2; ++gsite) {
                const int v = 2 * qr + b, u = 2 * k + (ng_x<N0>(b) ^ gsite);
                if (!in_a(u, v)) continue;
                float gd[3], go[3];
                const bool has = gd_at(u, v, gd, go);
                if (!gsite) {
                    const int c = ci(u, v);
                    const bool h = hz[c] != 0;
                    const float po = red_at(v) ? gd[2] : gd[0];
                    POh[c] = h ? po : 0.f;
                    POv[c] = h ? 0.f : po;
                }
                if (has && in_tile(u, v)) {
                    float d[3];
                    if (gsite) dg(u, v, b, d);
                    else dn(u, v, b, d);
                    param_grads(d, go, acc);
                }
            }
    }
    __syncthreads();
    // pass 2: green sites inside the area with all four neighbours in the extended tile
    for (int i = threadIdx.x; i < NC; i += kNT) {
        const int v = i / HWE, u = 2 * (i - v * HWE) + (ng_x<N0>(v & 1) ^ 1);
        if (u < 1 || u >= EW - 1 || v < 1 || v >= EH - 1 || !in_a(u, v)) continue;
        float gd[3], go[3];
        gd_at(u, v, gd, go);
        float ar = gd[0], ab = gd[2];
        // R neighbour: B = r + mean(B) - mean(R); B neighbour: R = b + mean(R) - mean(B)
        auto take = [&](int nu, int nv, const float* P) {
            if (!src(nu, nv)) return;
            const float w = 0.5f * P[ci(nu, nv)];
            if (red_at(nv)) { ar -= w; ab += w; }
            else { ar += w; ab -= w; }
        };
        virtuals(u, v, [&](int zu, int zv) {
            take(zu - 1, zv, POh);
            take(zu + 1, zv, POh);
            take(zu, zv - 1, POv);
            take(zu, zv + 1, POv);
        });
        AR[i] = ar;
        AB[i] = ab;
        if (in_tile(u, v)) RP[(v - 4) * TW + (u - 4)] = ar + ab + gd[1];
    }
    __syncthreads();
    // pass 3: non-green sites of tile +-2 inside the area
    for (int i = threadIdx.x; i < NC; i += kNT) {
        const int v = i / HWE, u = 2 * (i - v * HWE) + ng_x<N0>(v & 1);
        if (u < 2 || u >= EW - 2 || v < 2 || v >= EH - 2 || !in_a(u, v)) continue;
        const float* A = red_at(v) ? AR : AB;
        float S = 0.f;
        virtuals(u, v, [&](int zu, int zv) {
            if (src(zu - 1, zv)) S += A[ci(zu - 1, zv)];
            if (src(zu + 1, zv)) S += A[ci(zu + 1, zv)];
            if (src(zu, zv - 1)) S += A[ci(zu, zv - 1)];
            if (src(zu, zv + 1)) S += A[ci(zu, zv + 1)];
        });
        float gd[3], go[3];
        gd_at(u, v, gd, go);
        const float ag = gd[1] - 0.5f * S;
        const bool h = hz[i] != 0;
        AGh[i] = h ? ag : 0.f;
        AGv[i] = h ? 0.f : ag;
        if (in_tile(u, v)) RP[(v - 4) * TW + (u - 4)] = (gd[0] + gd[2]) + 0.5f * S;   // the own sample feeds both r and b
    }
    __syncthreads();
    // pass 4: raw adjoint of the tile's samples inside the area
    for (int i = threadIdx.x; i < TW * TH; i += kNT) {
        const int v = 4 + i / TW, u = 4 + (i - (i / TW) * TW);
        if (!in_a(u, v)) continue;
        const bool green = (u & 1) != ng_x<N0>(v & 1);
        const int step = green ? 1 : 2;            // directional-green taps reaching this sample
        const float w = green ? 0.5f : -0.25f;
        float r = RP[i];
        if (!green) {
            const int c = ci(u, v);
            r += 0.5f * (AGh[c] + AGv[c]);
        }
        int x1, x2, y1, y2;
        const int nxv = axis(ex + u, ax, aw, x1, x2), nyv = axis(ey + v, ay, ah, y1, y2);
        for (int iy = 0; iy < nyv; ++iy) {
            const int zv = (iy == 0 ? ey + v : (iy == 1 ? y1 : y2)) - ey;
            for (int ix = 0; ix < nxv; ++ix) {
                const int zu = (ix == 0 ? ex + u : (ix == 1 ? x1 : x2)) - ex;
                if (src(zu - step, zv)) r += w * AGh[ci(zu - step, zv)];
                if (src(zu + step, zv)) r += w * AGh[ci(zu + step, zv)];
                if (src(zu, zv - step)) r += w * AGv[ci(zu, zv - step)];
                if (src(zu, zv + step)) r += w * AGv[ci(zu, zv + step)];
            }
        }
        if (a.graw) a.graw[(size_t)(ty + v - 4 - ay) * aw + (size_t)(tx + u - 4 - ax)] = r;
    }
    flush_params(a, acc, red);
}

// Demosaic output of a non-green pixel (MenonDemosaic.cpp:162-173), as the
// forward takes it: red/blue means at the two neighbours along the site's
// decision (indices ia, ic of the minus / plus neighbour), green, own sample.
__device__ __forceinline__ void finish_nongreen(bool red_site, float self, double g, const double* red,
                                                const double* blue, int ia, int ic, float d[3]) {
    const double mr = 0.5 * (red[ia] + red[ic]);
    const double mb = 0.5 * (blue[ia] + blue[ic]);
    const float other = (float)((double)self + (red_site ? mb : mr) - (red_site ? mr : mb));
    d[0] = red_site ? self : other;
    d[1] = (float)g;
    d[2] = red_site ? other : self;
}

// Recomputing adjoint of one tile: the forward's row-pair stages over the
// extended tile (decisions included), then the gather passes.  LDS reuse
// (MenonFwdLds of the extended tile): POh/POv over dv (dead after the
// classifier); interior: AR/AB over gh, AGh/AGv over gv (dead after pass 1);
// border: AR/AB over dh, AGh/AGv over gv (blue and red, dead after pass 1),
// raw partials over gh; the parameter reduction over the raw codes.
template <int TW, int TH, int N0, bool BORDER>
__device__ __forceinline__ void menon_back_recompute(const BackwardArgs& a, const SensorArgs& s, char* lds, int tx,
                                                     int ty) {
    constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NC = HWE * EH;
    constexpr int J = (HWE * (EH / 2) + kNT - 1) / kNT;
    using F = MenonFwd<EW, EH, kNT, N0, BORDER>;
    using L = MenonFwdLds<EW, EH>;
    static_assert(2 * NC * 4 <= L::DVP * L::DVR * 8, "POh/POv do not fit over dv");
    static_assert(2 * NC * 4 <= L::NP * L::NR * 8 && 2 * NC * 4 <= L::DHP * L::DHR * 8, "AR/AB do not fit");
    static_assert(2 * NC * 4 <= L::NP * L::NR * 8 && TW * TH * 4 <= L::NP * L::NR * 8, "AG/raw partials do not fit");
    static_assert((kNT / 64) * 9 * 8 <= L::OFF_GH, "the parameter reduction does not fit over the codes");
    const int ex = tx - 4, ey = ty - 4;
    F t(s, lds, ex, ey);
    t.load();
    Upstream<J> up;
    if constexpr (!BORDER) fetch_upstream<EW, EH, N0>(a, ex, ey, up);   // in flight during the recompute
    __syncthreads();
    t.directional();
    __syncthreads();
    t.gradients();
    __syncthreads();
    t.classify();
    __syncthreads();
    t.chroma();
    __syncthreads();
    auto dn = [&](int u, int v, int b, float d[3]) {
        const bool h = t.hz[v * HWE + (u >> 1)] != 0;
        int xa = u - (h ? 1 : 0), ya = v - (h ? 0 : 1), xc = u + (h ? 1 : 0), yc = v + (h ? 0 : 1);
        t.mirror(xa, ya);
        t.mirror(xc, yc);
        finish_nongreen(((t.red_rows >> b) & 1u) != 0, t.Rn(F::rawi(u, v), b, F::nx(b)), t.gh[F::ni(u, v)], t.red,
                        t.blue, F::gi(xa, ya), F::gi(xc, yc), d);
    };
    auto dg = [&](int u, int v, int b, float d[3]) {
        const int g = F::gi(u, v);
        d[0] = (float)t.red[g];
        d[1] = t.Rn(F::rawi(u, v), b, F::nx(b) ^ 1);
        d[2] = (float)t.blue[g];
    };
    float* POh = reinterpret_cast<float*>(lds + L::OFF_DV);
    double* red = reinterpret_cast<double*>(lds + L::OFF_RAW);
    if constexpr (!BORDER) {
        float* AR = reinterpret_cast<float*>(lds + L::OFF_GH);
        float* AGh = reinterpret_cast<float*>(lds + L::OFF_GV);
        interior_gather<TW, TH, N0>(a, s, tx, ty, up, t.hz, t.red_rows, POh, POh + NC, AR, AR + NC, AGh, AGh + NC,
                                    red, dn, dg);
    } else {
        float* AR = reinterpret_cast<float*>(lds + L::OFF_DH);
        float* AGh = reinterpret_cast<float*>(lds + L::OFF_GV);
        float* RP = reinterpret_cast<float*>(lds + L::OFF_GH);
        border_gather<TW, TH, N0>(a, s, tx, ty, t.hz, t.red_rows, POh, POh + NC, AR, AR + NC, AGh, AGh + NC, RP,
                                  red, dn, dg);
    }
}

template <int TW, int TH>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(6))) menon_backward(const BackwardArgs args) {
    __shared__ __attribute__((aligned(16))) char lds[MenonFwdLds<TW + 8, TH + 8>::BYTES];
    // The arguments are read where they lie, in the kernel-argument segment
    // (the only parameter, at offset 0): a by-value copy that the optimiser
    // cannot fold away would be materialised in scratch memory.
    (void)args;
    const BackwardArgs& a = *(const BackwardArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    int tx, ty;
    tile_of_border_first(a, tx, ty, TW, TH);
    // Uniform across the workgroup; said so, so that everything derived from
    // the tile origin (CFA parities, per-site levels) stays scalar.
    tx = __builtin_amdgcn_readfirstlane(tx);
    ty = __builtin_amdgcn_readfirstlane(ty);
    const SensorArgs& s = a.s;
    const bool interior = tx - 10 >= s.ax && ty - 10 >= s.ay && tx + TW + 10 <= s.ax + s.aw &&
                          ty + TH + 10 <= s.ay + s.ah;
    // x parity of the non-green sites in even tile rows: the same for every
    // tile of the launch (tile origins step by even sizes)
    const uint32_t n0 = ((s.ng_parity >> ((uint32_t)ty & 1u)) & 1u) ^ ((uint32_t)tx & 1u);
    if (n0) {
        if (interior) menon_back_recompute<TW, TH, 1, false>(a, s, lds, tx, ty);
        else menon_back_recompute<TW, TH, 1, true>(a, s, lds, tx, ty);
    } else {
        if (interior) menon_back_recompute<TW, TH, 0, false>(a, s, lds, tx, ty);
        else menon_back_recompute<TW, TH, 0, true>(a, s, lds, tx, ty);
    }
}

// ------------------------------------------- Menon, recorded decisions --
//
// Tiles whose reach (tile +-4, within the active area) has decisions
// recorded by the forward (rawops_render_record): no classifier recompute.
// The selected green at the non-green sites of tile +-2 and red/blue at the
// green sites of tile +-1 (for the demosaic output of the tile's pixels) are
// recomputed from the samples of tile +-4 in the forward's FP64 operation
// order, reads mirrored at the area edge as the forward mirrors them.
template <int TW, int TH>
struct RecLds {
    static constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NC = HWE * EH;
    static constexpr int GW = (TW + 4) / 2, GH = TH + 4;                   // tile +-2, compacted
    static constexpr int OFF_RAW = 0;                                       // u16 codes, EW x EH
    static constexpr int OFF_HZ = (EW * EH * 2 + 15) / 16 * 16;             // decisions, NC
    static constexpr int OFF_GREEN = (OFF_HZ + NC + 15) / 16 * 16;          // f64 GW x GH
    static constexpr int OFF_RED = OFF_GREEN + GW * GH * 8;
    static constexpr int OFF_BLUE = OFF_RED + GW * GH * 8;
    static constexpr int OFF_PO = OFF_BLUE + GW * GH * 8;                   // f32 POh, POv
    static constexpr int BYTES = OFF_PO + 2 * NC * 4;
    // After pass 1: AR/AB over green + red, raw partials over blue; after
    // pass 2: AGh/AGv over PO; the parameter reduction over the codes.
    static_assert(2 * NC * 4 <= 2 * GW * GH * 8, "AR/AB do not fit over green/red");
    static_assert(TW * TH * 4 <= GW * GH * 8, "raw partials do not fit over blue");
    static_assert((kNT / 64) * 9 * 8 <= OFF_HZ, "the parameter reduction does not fit over the codes");
};

template <int TW, int TH, int N0, bool BORDER>
__device__ __forceinline__ void menon_back_recorded_tile(const BackwardArgs& a, const SensorArgs& s, char* lds,
                                                         int tx, int ty) {
    using R = RecLds<TW, TH>;
    constexpr int EW = R::EW, EH = R::EH, HWE = R::HWE, NC = R::NC, GW = R::GW;
    constexpr int J = (HWE * (EH / 2) + kNT - 1) / kNT;
    constexpr int NG = GW * ((EH - 4) / 2);                         // row-pair items over tile +-2
    const int ex = tx - 4, ey = ty - 4;
    uint16_t* raw = reinterpret_cast<uint16_t*>(lds + R::OFF_RAW);
    uint8_t* hz = reinterpret_cast<uint8_t*>(lds + R::OFF_HZ);
    double* green = reinterpret_cast<double*>(lds + R::OFF_GREEN);
    double* red = reinterpret_cast<double*>(lds + R::OFF_RED);
    double* blue = reinterpret_cast<double*>(lds + R::OFF_BLUE);
    SiteLevels lv;
    lv.init(s, ex, ey);
    uint32_t red_rows = 0;
    for (int b = 0; b < 2; ++b) red_rows |= (color_of(s, ex + ng_x<N0>(b), ey + b) == 0 ? 1u : 0u) << b;
    auto inside = [&](int u, int v) { return !BORDER || in_area(s, ex + u, ey + v); };
    auto mirror = [&](int& u, int& v) {
        if (BORDER) {
            const int sx = ex + u, sy = ey + v;
            if (!in_area(s, sx, sy)) { u = reflect(sx, s.ax, s.aw) - ex; v = reflect(sy, s.ay, s.ah) - ey; }
        }
    };
    // sample at (u + ox, v + oy), mirrored; parity of the unmirrored position (reflection keeps it)
    auto Rd = [&](int u, int v, int ox, int oy) -> double {
        const int py = (v + oy) & 1, px = (u + ox) & 1;
        int x = u + ox, y = v + oy;
        mirror(x, y);
        return (double)lv.norm(raw[y * EW + x], py, px);
    };
    auto g2 = [](int u, int v) { return (v - 2) * GW + ((u - 2) >> 1); };   // tile +-2 compaction
    auto G = [&](int u, int v, int ox, int oy) -> double {
        int x = u + ox, y = v + oy;
        mirror(x, y);
        return green[g2(x, y)];
    };

    // Codes of the extended tile and its recorded decisions: every load
    // issued (unconditionally, at clamped in-bounds addresses) before any is
    // consumed.  The upstream gradients are fetched after the LDS stores, so
    // they stay in flight through the recompute below.
    {
        constexpr int NW = EW * EH / 2, JW = (NW + kNT - 1) / kNT, JD = (NC + kNT - 1) / kNT;
        const int64_t stride = s.stride;
        const uint16_t* base = s.samples + ((int64_t)ey * stride + ex);
        const bool pairs = !BORDER && ((reinterpret_cast<uintptr_t>(base) | (uintptr_t)(stride * 2)) & 3) == 0;
        uint32_t w[JW], dec[JD];
        if (pairs) {
#pragma unroll
            for (int j = 0; j < JW; ++j) {
                const int i = min((int)threadIdx.x + j * kNT, NW - 1);
                const int v = i / (EW / 2), k = i - v * (EW / 2);
                // 32-bit offsets (24-bit products) from the uniform base
                w[j] = *reinterpret_cast<const uint32_t*>(base + (__umul24((uint32_t)v, (uint32_t)stride) + 2u * k));
            }
        }
#pragma unroll
        for (int j = 0; j < JD; ++j) {
            // the word holding the site's bit (clamped into the record: sites
            // outside the area are not read)
            const int i = min((int)threadIdx.x + j * kNT, NC - 1);
            const int v = i / HWE, u = 2 * (i - v * HWE) + ng_x<N0>(v & 1);
            const int sx = min(max(ex + u, a.rx), a.rx + a.rw - 1), sy = min(max(ey + v, a.ry), a.ry + a.rh - 1);
            const uint32_t X = (uint32_t)(sx - a.rx), Y = (uint32_t)(sy - a.ry);
            dec[j] = a.decisions[((__umul24(Y >> 5, a.dtiles) + (X >> 6)) << 5) + (Y & 31u)] >> ((X >> 1) & 31u);
        }
        if (pairs) {
#pragma unroll
            for (int j = 0; j < JW; ++j) {
                const int i = (int)threadIdx.x + j * kNT;
                const int v = i / (EW / 2), k = i - v * (EW / 2);
                if (i < NW) *reinterpret_cast<uint32_t*>(raw + v * EW + 2 * k) = w[j];
            }
        } else {
            for (int i = threadIdx.x; i < NW; i += kNT) {
                const int v = i / (EW / 2), k = i - v * (EW / 2);
                const uint16_t* p = base + (int64_t)v * stride + 2 * k;
                if (inside(2 * k, v)) raw[v * EW + 2 * k] = p[0];
                if (inside(2 * k + 1, v)) raw[v * EW + 2 * k + 1] = p[1];
            }
        }
#pragma unroll
        for (int j = 0; j < JD; ++j) {
            const int i = (int)threadIdx.x + j * kNT;
            const int v = i / HWE, u = 2 * (i - v * HWE) + ng_x<N0>(v & 1);
            if (i < NC && inside(u, v)) hz[i] = (uint8_t)(dec[j] & 1u);
        }
    }
    Upstream<J> up;
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (!BORDER) fetch_upstream<EW, EH, N0>(a, ex, ey, up);
    __syncthreads();
    // selected green at non-green sites of tile +-2 (MenonDemosaic.cpp:107-119, 141-147)
    for (int it = threadIdx.x; it < NG; it += kNT) {
        const int qr = it / GW, k = it - qr * GW;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 + 2 * qr + b, u = 2 + 2 * k + ng_x<N0>(b);
            if (!inside(u, v)) continue;
            const double own = Rd(u, v, 0, 0);
            double gsel;
            if (hz[v * HWE + (u >> 1)])
                gsel = 0.5 * (Rd(u, v, -1, 0) + Rd(u, v, 1, 0)) + 0.5 * own - 0.25 * (Rd(u, v, -2, 0) + Rd(u, v, 2, 0));
            else
                gsel = 0.5 * (Rd(u, v, 0, -1) + Rd(u, v, 0, 1)) + 0.5 * own - 0.25 * (Rd(u, v, 0, -2) + Rd(u, v, 0, 2));
            green[g2(u, v)] = gsel;
        }
    }
    __syncthreads();
    // red and blue at green sites of tile +-1 (MenonDemosaic.cpp:148-161)
    for (int it = threadIdx.x; it < NG; it += kNT) {
        const int qr = it / GW, k = it - qr * GW;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 + 2 * qr + b, u = 2 + 2 * k + (ng_x<N0>(b) ^ 1);
            if (u < 3 || u >= EW - 3 || v < 3 || v >= EH - 3 || !inside(u, v)) continue;
            const double own = Rd(u, v, 0, 0);
            const double ch = own + 0.5 * (Rd(u, v, -1, 0) + Rd(u, v, 1, 0)) - 0.5 * (G(u, v, -1, 0) + G(u, v, 1, 0));
            const double cv = own + 0.5 * (Rd(u, v, 0, -1) + Rd(u, v, 0, 1)) - 0.5 * (G(u, v, 0, -1) + G(u, v, 0, 1));
            const bool red_row = (red_rows >> b) & 1u;
            red[g2(u, v)] = red_row ? ch : cv;
            blue[g2(u, v)] = red_row ? cv : ch;
        }
    }
    __syncthreads();
    auto dn = [&](int u, int v, int b, float d[3]) {
        const bool h = hz[v * HWE + (u >> 1)] != 0;
        int xa = u - (h ? 1 : 0), ya = v - (h ? 0 : 1), xc = u + (h ? 1 : 0), yc = v + (h ? 0 : 1);
        mirror(xa, ya);
        mirror(xc, yc);
        finish_nongreen(((red_rows >> b) & 1u) != 0, lv.norm(raw[v * EW + u], b, ng_x<N0>(b)), green[g2(u, v)], red,
                        blue, g2(xa, ya), g2(xc, yc), d);
    };
    auto dg = [&](int u, int v, int b, float d[3]) {
        const int g = g2(u, v);
        d[0] = (float)red[g];
        d[1] = lv.norm(raw[v * EW + u], b, ng_x<N0>(b) ^ 1);
        d[2] = (float)blue[g];
    };
    float* POh = reinterpret_cast<float*>(lds + R::OFF_PO);
    float* AR = reinterpret_cast<float*>(lds + R::OFF_GREEN);
    double* rbuf = reinterpret_cast<double*>(lds + R::OFF_RAW);
    if constexpr (!BORDER)
        interior_gather<TW, TH, N0>(a, s, tx, ty, up, hz, red_rows, POh, POh + NC, AR, AR + NC, POh, POh + NC, rbuf,
                                    dn, dg);
    else
        border_gather<TW, TH, N0>(a, s, tx, ty, hz, red_rows, POh, POh + NC, AR, AR + NC, POh, POh + NC,
                                  reinterpret_cast<float*>(lds + R::OFF_BLUE), rbuf, dn, dg);
}

// One workgroup per tile of the recorded block [r0, r1) x [c0, c1), its
// mirrored (slower) tiles first so they do not trail the launch.
template <int TW, int TH>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(8))) menon_backward_recorded(const BackwardArgs args) {
    __shared__ __attribute__((aligned(16))) char lds[RecLds<TW, TH>::BYTES];
    (void)args;
    const BackwardArgs& a = *(const BackwardArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    int r, c;
    border_first((int)blockIdx.x, a.c1 - a.c0, a.r1 - a.r0, a.ir0, a.ir1, a.ic0, a.ic1, r, c);
    r += a.r0;
    c += a.c0;
    const int tx = __builtin_amdgcn_readfirstlane(a.s.ax + (a.tile0_x + c) * TW);
    const int ty = __builtin_amdgcn_readfirstlane(a.s.ay + (a.tile0_y + r) * TH);
    const SensorArgs& s = a.s;
    // no read of tile +-4 leaves the area: no mirroring
    const bool interior = tx - 4 >= s.ax && ty - 4 >= s.ay && tx + TW + 4 <= s.ax + s.aw && ty + TH + 4 <= s.ay + s.ah;
    const uint32_t n0 = ((s.ng_parity >> ((uint32_t)ty & 1u)) & 1u) ^ ((uint32_t)tx & 1u);
    if (n0) {
        if (interior) menon_back_recorded_tile<TW, TH, 1, false>(a, s, lds, tx, ty);
        else menon_back_recorded_tile<TW, TH, 1, true>(a, s, lds, tx, ty);
    } else {
        if (interior) menon_back_recorded_tile<TW, TH, 0, false>(a, s, lds, tx, ty);
        else menon_back_recorded_tile<TW, TH, 0, true>(a, s, lds, tx, ty);
    }
}

// ------------------------------------------------------------- bilinear --

template <int TW, int TH>
__global__ void __launch_bounds__(kNT) bilinear_backward(const BackwardArgs a) {
    __shared__ float raw[(TW + 2) * (TH + 2)];
    __shared__ float adj[TW * TH];
    __shared__ double red[(kNT / 64) * 9];
    int tx, ty;
    tile_of(a, tx, ty, TW, TH);
    const SensorArgs& s = a.s;
    load_raw<TW, TH, 1, true, kNT>(s, tx, ty, raw);
    for (int i = threadIdx.x; i < TW * TH; i += kNT) adj[i] = 0.0f;
    __syncthreads();
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int EW = TW + 2, EH = TH + 2;
    for (int i = threadIdx.x; i < EW * EH; i += kNT) {
        const int v = i / EW, u = i - v * EW;       // tile coordinates + 1
        const int sx = tx + u - 1, sy = ty + v - 1;
        if (!in_area(s, sx, sy)) continue;
        float go[3];
        if (!load_gout(a, sx, sy, go)) continue;
        float gd[3];
        chain_back(a, go, gd);
        const bool mine = u >= 1 && u <= TW && v >= 1 && v <= TH;
        if (mine) {
            float d[3];
            bilinear_pixel<TW, TH, true>(s, raw, tx, ty, u - 1, v - 1, d[0], d[1], d[2]);
            param_grads(d, go, acc);
        }
        const int own = color_of(s, sx, sy);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (own == c) {
                if (mine) atomicAdd(adj + (v - 1) * TW + (u - 1), gd[c]);
                continue;
            }
            int count = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    if (in_area(s, sx + dx, sy + dy) && color_of(s, sx + dx, sy + dy) == c) ++count;
            if (!count) continue;
            const float w = gd[c] / (float)count;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int px = u - 1 + dx, py = v - 1 + dy;
                    if (px < 0 || px >= TW || py < 0 || py >= TH) continue;
                    if (!in_area(s, sx + dx, sy + dy) || color_of(s, sx + dx, sy + dy) != c) continue;
                    atomicAdd(adj + py * TW + px, w);
                }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TW * TH; i += kNT) {
        const int v = i / TW, u = i - v * TW;
        const int sx = tx + u, sy = ty + v;
        if (a.graw && sx < s.ax + s.aw && sy < s.ay + s.ah)
            a.graw[(size_t)(sy - s.ay) * s.aw + (size_t)(sx - s.ax)] = adj[i];
    }
    flush_params(a, acc, red);
}

}  // namespace

hipError_t launch_render_backward(const BackwardLaunch& L, hipStream_t stream) {
    BackwardArgs a{};
    a.s = L.sensor;
    a.vx = (int)L.viewport.x; a.vy = (int)L.viewport.y;
    a.vw = (int)L.viewport.width; a.vh = (int)L.viewport.height;
    for (int i = 0; i < 3; ++i) a.g[i] = L.gains[i];
    a.m = L.multiplier;
    for (int i = 0; i < 9; ++i) a.M[i] = L.matrix[i];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            a.K[3 * c + r] = (float)((double)L.gains[c] * (double)L.multiplier * (double)L.matrix[3 * r + c]);
    a.gout = L.g_out;
    a.gstride = L.g_stride;
    a.graw = L.g_raw;
    a.gparams = L.g_params;
    // Only raw samples within 4 of the viewport receive gradient (the
    // adjoint's reach); tiles are anchored at the active-area origin, 64 x
    // kAdjTH (Menon: the recorded kernel 48 KB of LDS, three workgroups per
    // CU; the recomputing kernel, whose recompute region is tile +-4, 71 KB,
    // two) or 64 x 32 (bilinear).
    const bool menon = L.demosaic == RAWOPS_DEMOSAIC_MENON;
    const int TW = 64, TH = menon ? kAdjTH : 32;
    const int x0 = std::max(0, a.vx - 4 - a.s.ax), y0 = std::max(0, a.vy - 4 - a.s.ay);
    const int x1 = std::min(a.s.aw, a.vx + a.vw + 4 - a.s.ax), y1 = std::min(a.s.ah, a.vy + a.vh + 4 - a.s.ay);
    a.tile0_x = x0 / TW;
    a.tile0_y = y0 / TH;
    a.tiles_x = (x1 + TW - 1) / TW - a.tile0_x;
    const int tiles_y = (y1 + TH - 1) / TH - a.tile0_y;
    a.tiles_y = tiles_y;
    if (a.vw <= 0 || a.vh <= 0 || a.tiles_x <= 0 || tiles_y <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)a.tiles_x * (unsigned)tiles_y;
    // Interior block of the Menon launch: tiles whose recompute support
    // (tile +-10) lies inside the active area, a contiguous range of rows
    // and of columns; empty when either range is.
    auto interior_range = [](int first, int n, int T, int lo, int len, int& i0, int& i1) {
        i0 = n; i1 = n;
        for (int i = 0; i < n; ++i) {
            const int t0 = lo + (first + i) * T;
            if (t0 - 10 >= lo && t0 + T + 10 <= lo + len) { if (i0 == n) i0 = i; i1 = i + 1; }
        }
        if (i0 >= i1) i0 = i1 = 0;
    };
    interior_range(a.tile0_y, tiles_y, TH, a.s.ay, a.s.ah, a.r0, a.r1);
    interior_range(a.tile0_x, a.tiles_x, TW, a.s.ax, a.s.aw, a.c0, a.c1);
    if (menon && L.decisions) {
        // Tiles whose reach within the area ((tile +-4) n area) lies inside
        // the recorded viewport read the recorded decisions: per axis a
        // contiguous range of the launch's tiles.  The others recompute.
        auto covered = [](int first, int n, int T, int lo, int len, int rec_lo, int rec_len, int& i0, int& i1) {
            i0 = n; i1 = 0;
            for (int i = 0; i < n; ++i) {
                const int t0 = lo + (first + i) * T;
                const int r0 = std::max(t0 - 4, lo), r1 = std::min(t0 + T + 4, lo + len);
                if (r0 >= rec_lo && r1 <= rec_lo + rec_len) { i0 = std::min(i0, i); i1 = std::max(i1, i + 1); }
            }
            if (i0 >= i1) i0 = i1 = 0;
        };
        covered(a.tile0_y, tiles_y, TH, a.s.ay, a.s.ah, (int)L.recorded.y, (int)L.recorded.height, a.r0, a.r1);
        covered(a.tile0_x, a.tiles_x, TW, a.s.ax, a.s.aw, (int)L.recorded.x, (int)L.recorded.width, a.c0, a.c1);
        if (a.r0 == a.r1 || a.c0 == a.c1) a.r0 = a.r1 = a.c0 = a.c1 = 0;
        // within it, the tiles whose reach stays inside the area (no mirroring)
        auto unmirrored = [](int first, int from, int to, int T, int lo, int len, int& i0, int& i1) {
            i0 = to; i1 = from;
            for (int i = from; i < to; ++i) {
                const int t0 = lo + (first + i) * T;
                if (t0 - 4 >= lo && t0 + T + 4 <= lo + len) { i0 = std::min(i0, i); i1 = std::max(i1, i + 1); }
            }
            if (i0 >= i1) i0 = i1 = from;
            i0 -= from; i1 -= from;
        };
        unmirrored(a.tile0_y, a.r0, a.r1, TH, a.s.ay, a.s.ah, a.ir0, a.ir1);
        unmirrored(a.tile0_x, a.c0, a.c1, TW, a.s.ax, a.s.aw, a.ic0, a.ic1);
        if (a.ir0 == a.ir1 || a.ic0 == a.ic1) a.ir0 = a.ir1 = a.ic0 = a.ic1 = 0;
        a.decisions = L.decisions;
        a.dtiles = decision_tiles(L.recorded.width);
        a.rx = (int)L.recorded.x; a.ry = (int)L.recorded.y;
        a.rw = (int)L.recorded.width; a.rh = (int)L.recorded.height;
        const unsigned rec = (unsigned)((a.r1 - a.r0) * (a.c1 - a.c0));
        if (rec) {
            hipLaunchKernelGGL((menon_backward_recorded<64, kAdjTH>), dim3(rec), dim3(kNT), 0, stream, a);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        // the recomputing kernel's border-first order lists the tiles outside
        // the block [r0, r1) x [c0, c1) first: launch only those
        if (blocks - rec)
            hipLaunchKernelGGL((menon_backward<64, kAdjTH>), dim3(blocks - rec), dim3(kNT), 0, stream, a);
        return hipGetLastError();
    }
    if (a.r0 == a.r1 || a.c0 == a.c1) a.r0 = a.r1 = a.c0 = a.c1 = 0;
    if (menon)
        hipLaunchKernelGGL((menon_backward<64, kAdjTH>), dim3(blocks), dim3(kNT), 0, stream, a);
    else
        hipLaunchKernelGGL((bilinear_backward<64, 32>), dim3(blocks), dim3(kNT), 0, stream, a);
    return hipGetLastError();
}

}  // namespace rawops
