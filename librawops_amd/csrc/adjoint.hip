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
#include "menon_tile.hpp"
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kNT = 512;

struct BackwardArgs {
    SensorArgs s;
    int vx, vy, vw, vh;        // viewport (native sensor coordinates)
    float g[3], m, M[9];
    const float* gout;         // viewport RGB
    uint32_t gstride;          // pixels per gout row
    float* graw;               // active-area plane, aw x ah
    double* gparams;           // 13: gains[3], multiplier, matrix[9] (row-major)
    int tiles_x, tiles_y;      // tiles of the launch
    int tile0_x, tile0_y;      // first tile (in tiles from the area origin)
    int r0, r1, c0, c1;        // launch-relative tile rows / columns of the interior block
    const uint8_t* decisions;  // recorded Menon decisions (area plane, dpitch bytes a row) or null
    uint32_t dpitch;
};

// dL/dout at a sensor pixel (zero outside the viewport).
__device__ __forceinline__ bool load_gout(const BackwardArgs& a, int sx, int sy, float go[3]) {
    const int lx = sx - a.vx, ly = sy - a.vy;
    if (lx < 0 || ly < 0 || lx >= a.vw || ly >= a.vh) return false;
    const float* p = a.gout + ((size_t)ly * a.gstride + (size_t)lx) * 3;
    go[0] = p[0]; go[1] = p[1]; go[2] = p[2];
    return true;
}

// gd_c = g_c * (m * (M^T gout)_c): gradient at the demosaic output.
__device__ __forceinline__ void chain_back(const BackwardArgs& a, const float go[3], float u[3], float gd[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        u[c] = a.M[c] * go[0] + a.M[3 + c] * go[1] + a.M[6 + c] * go[2];
        gd[c] = a.g[c] * (a.m * u[c]);
    }
}

// Parameter gradients.  With u = M^T gout and out = M (m g (.) d), every
// parameter gradient is a fixed combination of the nine sums
//     P[r][c] = sum over pixels of gout_r d_c :
//     dL/dg_c    = m sum_r M[r][c] P[r][c]
//     dL/dm      = sum_c g_c sum_r M[r][c] P[r][c]
//     dL/dM[r][c] = m g_c P[r][c]
// so a pixel adds nine exact FP64 products and the combinations are taken
// once per workgroup.
__device__ __forceinline__ void param_grads(const float d[3], const float go[3], double (&acc)[9]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r * 3 + c] += (double)go[r] * (double)d[c];
}

__device__ __forceinline__ void flush_params(const BackwardArgs& a, double (&acc)[9], double* red /* [waves][9] LDS */) {
#pragma unroll
    for (int i = 0; i < 9; ++i)
        for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int i = 0; i < 9; ++i) red[wave * 9 + i] = acc[i];
    __syncthreads();
    if (threadIdx.x < 13) {
        double P[9];
        for (int i = 0; i < 9; ++i) {
            double t = 0;
            for (int w = 0; w < kNT / 64; ++w) t += red[w * 9 + i];
            P[i] = t;
        }
        const int k = threadIdx.x;
        double v;
        if (k < 3) {
            v = (double)a.m * ((double)a.M[k] * P[k] + (double)a.M[3 + k] * P[3 + k] + (double)a.M[6 + k] * P[6 + k]);
        } else if (k == 3) {
            v = 0;
            for (int c = 0; c < 3; ++c)
                v += (double)a.g[c] * ((double)a.M[c] * P[c] + (double)a.M[3 + c] * P[3 + c] + (double)a.M[6 + c] * P[6 + c]);
        } else {
            const int c = (k - 4) % 3;
            v = (double)a.m * (double)a.g[c] * P[k - 4];
        }
        atomicAdd(a.gparams + k, v);
    }
}

// Border-first tile order: the tiles outside the interior block [r0, r1) x
// [c0, c1) (slower: mirrored reads, scatter adjoint) take the first block
// indices so they start with the launch instead of trailing it.
__device__ __forceinline__ void tile_of_border_first(const BackwardArgs& a, int& tx, int& ty, int TW, int TH) {
    const int TX = a.tiles_x, TY = a.tiles_y;
    const int top = a.r0 * TX, bottom = (TY - a.r1) * TX, side = a.c0 + (TX - a.c1), middle = (a.r1 - a.r0) * side;
    int b = (int)blockIdx.x, r, c;
    if (b < top) {
        r = b / TX; c = b - r * TX;
    } else if ((b -= top) < bottom) {
        r = a.r1 + b / TX; c = b - (b / TX) * TX;
    } else if ((b -= bottom) < middle) {
        r = a.r0 + b / side;
        const int j = b - (b / side) * side;
        c = j < a.c0 ? j : a.c1 + (j - a.c0);
    } else {
        b -= middle;
        const int ci = a.c1 - a.c0;
        r = a.r0 + b / ci; c = a.c0 + (b - (b / ci) * ci);
    }
    tx = a.s.ax + (a.tile0_x + c) * TW;
    ty = a.s.ay + (a.tile0_y + r) * TH;
}

__device__ __forceinline__ void tile_of(const BackwardArgs& a, int& tx, int& ty, int TW, int TH) {
    tx = a.s.ax + (a.tile0_x + (int)(blockIdx.x % a.tiles_x)) * TW;
    ty = a.s.ay + (a.tile0_y + (int)(blockIdx.x / a.tiles_x)) * TH;
}

// --------------------------------------------------------------- Menon --

// Interior tiles (forward support of tile +-4 inside the active area, so no
// read is mirrored): the forward is recomputed over the extended tile
// (tile +-4) with the row-pair forward stages, then the same transposes as
// the scatter path below are taken as gathers -- each stage's adjoint at a
// site is a fixed stencil over its neighbours and their recorded decisions --
// so no LDS atomics are needed.  Extended-tile coordinates (u, v) = sensor -
// (ex, ey); the tile is [4, 4+TW) x [4, 4+TH).
//
// Work is handed out as row-pair items like the forward stages: item (qr, k)
// holds, in each row v = 2 qr + b, the non-green site 2k + nx(b) and the
// green site 2k + 1 - nx(b).  Both have compacted index c = v (EW/2) + k in
// the adjoint planes, so every neighbour is a compile-time offset from c, and
// a thread keeps its items' per-site values in registers across passes:
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
// LDS reuse: POh/POv over dv (dead after the classifier), AR/AB over gh and
// AGh/AGv over gv (both dead after pass 1), the parameter reduction over raw.
template <int TW, int TH, int N0>
__device__ __forceinline__ void menon_back_interior(const BackwardArgs& a, const SensorArgs& s, char* lds, int tx,
                                                    int ty) {
    constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NC = HWE * EH;
    constexpr int NI = HWE * (EH / 2), J = (NI + kNT - 1) / kNT;   // row-pair items
    using F = MenonFwd<EW, EH, kNT, N0, false>;
    using L = MenonFwdLds<EW, EH>;
    static_assert(2 * NC * 4 <= L::DVP * L::DVR * 8, "POh/POv do not fit over dv");
    static_assert(2 * NC * 4 <= L::NP * L::NR * 8, "AR/AB do not fit over gh");
    static_assert(2 * NC * 4 <= L::NP * L::NR * 8, "AGh/AGv do not fit over gv");
    static_assert((kNT / 64) * 9 * 8 <= L::OFF_GH, "the parameter reduction does not fit over raw");
    static_assert(TW % 2 == 0 && TH % 2 == 0, "tiles must cover whole Bayer cells");
    const int ex = tx - 4, ey = ty - 4;
    F t(s, lds, ex, ey);
    t.load();
    // Upstream gradients of the thread's sites, fetched while the forward is
    // recomputed.
    float goN[J][2][3], goG[J][2][3];
    uint32_t hasN = 0, hasG = 0;    // bit 2j + b
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it < NI) {
            const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int v = 2 * qr + b, un = 2 * k + F::nx(b);
                if (load_gout(a, ex + un, ey + v, goN[j][b])) hasN |= 1u << (2 * j + b);
                if (load_gout(a, ex + (un ^ 1), ey + v, goG[j][b])) hasG |= 1u << (2 * j + b);
            }
        }
    }
    __syncthreads();
    t.directional();
    __syncthreads();
    t.gradients();
    __syncthreads();
    t.classify();
    __syncthreads();
    t.chroma();
    __syncthreads();

    float* POh = reinterpret_cast<float*>(lds + L::OFF_DV);
    float* POv = POh + NC;
    float* AR = reinterpret_cast<float*>(lds + L::OFF_GH);
    float* AB = AR + NC;
    float* AGh = reinterpret_cast<float*>(lds + L::OFF_GV);
    float* AGv = AGh + NC;
    const bool red0 = (t.red_rows & 1u) != 0, red1 = (t.red_rows & 2u) != 0;   // non-green sites of row parity b are red
    auto red_row = [&](int b) { return b ? red1 : red0; };
    auto tile_item = [](int qr, int k) { return qr >= 2 && qr < 2 + TH / 2 && k >= 2 && k < 2 + TW / 2; };
    // per item and row: green-channel gd and gd_r + gd_b of the non-green
    // site, gd of the green site, raw partials of both
    float pg[J][2], rn[J][2], gg[J][2][3], rpg[J][2], rpn[J][2];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

    // pass 1
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
        const bool mine = tile_item(qr, k);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, nb = F::nx(b), un = 2 * k + nb, ug = un ^ 1, c = v * HWE + k;
            const bool rs = red_row(b);
            const bool hn = (hasN >> (2 * j + b)) & 1u, hg = (hasG >> (2 * j + b)) & 1u;
            float gd[3], uu[3];
            if (hn) chain_back(a, goN[j][b], uu, gd);
            else gd[0] = gd[1] = gd[2] = 0.f;
            const bool h = t.hz[c] != 0;
            const float po = rs ? gd[2] : gd[0];
            POh[c] = h ? po : 0.f;
            POv[c] = h ? 0.f : po;
            pg[j][b] = gd[1];
            rn[j][b] = gd[0] + gd[2];           // the own sample feeds both r and b
            if (hg) chain_back(a, goG[j][b], uu, gg[j][b]);
            else gg[j][b][0] = gg[j][b][1] = gg[j][b][2] = 0.f;
            if (mine) {
                if (hn) {
                    // final chroma (MenonDemosaic.cpp:162-173), as the forward takes it
                    const int ia = F::gi(un - (h ? 1 : 0), v - (h ? 0 : 1));
                    const int ic = F::gi(un + (h ? 1 : 0), v + (h ? 0 : 1));
                    const double mr = 0.5 * (t.red[ia] + t.red[ic]);
                    const double mb = 0.5 * (t.blue[ia] + t.blue[ic]);
                    const float self = t.Rn(F::rawi(un, v), b, nb);
                    const float other = (float)((double)self + (rs ? mb : mr) - (rs ? mr : mb));
                    const float d[3] = {rs ? self : other, (float)t.gh[F::ni(un, v)], rs ? other : self};
                    param_grads(d, goN[j][b], acc);
                }
                if (hg) {
                    const int g = F::gi(ug, v);
                    const float d[3] = {(float)t.red[g], t.Rn(F::rawi(ug, v), b, nb ^ 1), (float)t.blue[g]};
                    param_grads(d, goG[j][b], acc);
                }
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
            const int v = 2 * qr + b, nb = F::nx(b), ug = 2 * k + 1 - nb, c = v * HWE + k;
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
            const int v = 2 * qr + b, nb = F::nx(b), un = 2 * k + nb, c = v * HWE + k;
            if (un < 2 || un >= EW - 2 || v < 2 || v >= EH - 2) continue;
            const float* A = red_row(b) ? AR : AB;
            const float S = (A[c - 1 + nb] + A[c + nb]) + (A[c - HWE] + A[c + HWE]);
            const float ag = pg[j][b] - 0.5f * S;
            const bool h = t.hz[c] != 0;
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
            const int v = 2 * qr + b, nb = F::nx(b), un = 2 * k + nb, ug = un ^ 1, c = v * HWE + k;
            const float rgv = rpg[j][b] + 0.5f * (AGh[c - nb] + AGh[c + 1 - nb]) + 0.5f * (AGv[c - HWE] + AGv[c + HWE]);
            const float rnv = rpn[j][b] + 0.5f * (AGh[c] + AGv[c]) - 0.25f * (AGh[c - 1] + AGh[c + 1]) -
                              0.25f * (AGv[c - 2 * HWE] + AGv[c + 2 * HWE]);
            if (a.graw) {
                float* row = a.graw + (size_t)(ty + v - 4 - s.ay) * s.aw + (size_t)(tx - 4 - s.ax);
                row[un] = rnv;
                row[ug] = rgv;
            }
        }
    }
    flush_params(a, acc, reinterpret_cast<double*>(lds + L::OFF_RAW));
}

// Border tiles: some reads of the forward are mirrored at the active-area
// edge (whole-sample reflection m, applied per coordinate).  A stage output
// at p reads its input at m(p + k) for each tap k, so the adjoint reaching a
// target q gathers over the virtual positions V(q) = {z : m(z) = q} that lie
// within a tap's reach (2) of the area -- q itself and its reflections
// across nearby edges -- and applies the interior stencil at each z, counting
// only sources inside the area.  The forward is recomputed with the border
// form of the row-pair stages; the passes are those of the interior path,
// per site, with upstream gradients re-read where a pass needs them.
// LDS reuse: POh/POv over dv (dead after the classifier), AR/AB over dh,
// AGh/AGv over gv (blue and red, dead after pass 1), the tile's raw partials
// over gh, the parameter reduction over raw.
template <int TW, int TH, int N0>
__device__ __forceinline__ void menon_back_border(const BackwardArgs& a, const SensorArgs& s, char* lds, int tx, int ty) {
    constexpr int EW = TW + 8, EH = TH + 8, HWE = EW / 2, NC = HWE * EH;
    using F = MenonFwd<EW, EH, kNT, N0, true>;
    using L = MenonFwdLds<EW, EH>;
    static_assert(2 * NC * 4 <= L::DVP * L::DVR * 8, "POh/POv do not fit over dv");
    static_assert(2 * NC * 4 <= L::DHP * L::DHR * 8, "AR/AB do not fit over dh");
    static_assert(2 * NC * 4 <= L::NP * L::NR * 8, "AGh/AGv do not fit over gv");
    static_assert(TW * TH * 4 <= L::NP * L::NR * 8, "raw partials do not fit over gh");
    static_assert((kNT / 64) * 9 * 8 <= L::OFF_GH, "the parameter reduction does not fit over raw");
    const int ex = tx - 4, ey = ty - 4;     // origin of the extended tile (tile +-4)
    F t(s, lds, ex, ey);
    t.load();
    __syncthreads();
    t.directional();
    __syncthreads();
    t.gradients();
    __syncthreads();
    t.classify();
    __syncthreads();
    t.chroma();
    __syncthreads();

    float* POh = reinterpret_cast<float*>(lds + L::OFF_DV);    // dv is dead after the classifier
    float* POv = POh + NC;
    float* AR = reinterpret_cast<float*>(lds + L::OFF_DH);     // dh (blue) is dead after pass 1
    float* AB = AR + NC;
    float* AGh = reinterpret_cast<float*>(lds + L::OFF_GV);
    float* AGv = AGh + NC;
    float* RP = reinterpret_cast<float*>(lds + L::OFF_GH);
    auto ci = [](int u, int v) { return v * HWE + (u >> 1); };
    auto in_tile = [](int u, int v) { return u >= 4 && u < 4 + TW && v >= 4 && v < 4 + TH; };
    auto in_ext = [](int u, int v) { return u >= 0 && u < EW && v >= 0 && v < EH; };
    // Area bounds read once (the kernel arguments stay in the argument
    // segment instead of being copied to scratch memory).
    const int ax = s.ax, ay = s.ay, aw = s.aw, ah = s.ah;
    auto in_a = [&](int u, int v) {
        const int x = ex + u, y = ey + v;
        return x >= ax && y >= ay && x < ax + aw && y < ay + ah;
    };
    auto src = [&](int u, int v) { return in_ext(u, v) && in_a(u, v); };    // a computed stage output
    auto red_at = [&](int, int v) { return ((t.red_rows >> (v & 1)) & 1u) != 0; };   // non-green sites only
    auto gd_at = [&](int u, int v, float gd[3], float go[3]) -> bool {
        float uu[3];
        if (!load_gout(a, ex + u, ey + v, go)) { gd[0] = gd[1] = gd[2] = 0.f; return false; }
        chain_back(a, go, uu, gd);
        return true;
    };
    // Virtual positions of (u, v): per coordinate, the position itself and
    // the outside positions within 2 of the area that reflect onto it (at
    // most two; kept in scalars, not an indexed array, so nothing goes to
    // scratch memory).
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
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

    // pass 1: every site of the extended tile inside the area, as row-pair
    // items (compile-time CFA parities: the per-site levels are then selected
    // at compile time)
    for (int it = threadIdx.x; it < NC / 2; it += kNT) {
        const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int gsite = 0; gsite < 2; ++gsite) {
                const int v = 2 * qr + b, pu = F::nx(b) ^ gsite, u = 2 * k + pu;   // pu: parity of u
                if (!in_a(u, v)) continue;
                float gd[3], go[3];
                const bool has = gd_at(u, v, gd, go);
                const bool rs = ((t.red_rows >> b) & 1u) != 0;
                if (!gsite) {
                    const int c = ci(u, v);
                    const bool h = t.hz[c] != 0;
                    const float po = rs ? gd[2] : gd[0];
                    POh[c] = h ? po : 0.f;
                    POv[c] = h ? 0.f : po;
                }
                if (has && in_tile(u, v)) {
                    // demosaic output (MenonDemosaic.cpp:162-173), as the forward takes it
                    float d[3];
                    if (gsite) {
                        const int g = F::gi(u, v);
                        d[0] = (float)t.red[g];
                        d[1] = t.Rn(F::rawi(u, v), b, pu);
                        d[2] = (float)t.blue[g];
                    } else {
                        const bool h = t.hz[ci(u, v)] != 0;
                        int xa = u - (h ? 1 : 0), ya = v - (h ? 0 : 1), xc = u + (h ? 1 : 0), yc = v + (h ? 0 : 1);
                        t.mirror(xa, ya);
                        t.mirror(xc, yc);
                        const int ia = F::gi(xa, ya), ic = F::gi(xc, yc);
                        const double mr = 0.5 * (t.red[ia] + t.red[ic]);
                        const double mb = 0.5 * (t.blue[ia] + t.blue[ic]);
                        const float self = t.Rn(F::rawi(u, v), b, pu);
                        const float other = (float)((double)self + (rs ? mb : mr) - (rs ? mr : mb));
                        d[0] = rs ? self : other;
                        d[1] = (float)t.gh[F::ni(u, v)];
                        d[2] = rs ? other : self;
                    }
                    param_grads(d, go, acc);
                }
            }
    }
    __syncthreads();
    // pass 2: green sites inside the area with all four neighbours in the extended tile
    for (int i = threadIdx.x; i < NC; i += kNT) {
        const int v = i / HWE, u = 2 * (i - v * HWE) + (F::nx(v & 1) ^ 1);
        if (u < 1 || u >= EW - 1 || v < 1 || v >= EH - 1 || !in_a(u, v)) continue;
        float gd[3], go[3];
        gd_at(u, v, gd, go);
        float ar = gd[0], ab = gd[2];
        // R neighbour: B = r + mean(B) - mean(R); B neighbour: R = b + mean(R) - mean(B)
        auto take = [&](int nu, int nv, const float* P) {
            if (!src(nu, nv)) return;
            const float w = 0.5f * P[ci(nu, nv)];
            if (red_at(nu, nv)) { ar -= w; ab += w; }
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
        const int v = i / HWE, u = 2 * (i - v * HWE) + F::nx(v & 1);
        if (u < 2 || u >= EW - 2 || v < 2 || v >= EH - 2 || !in_a(u, v)) continue;
        const float* A = red_at(u, v) ? AR : AB;
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
        const bool h = t.hz[i] != 0;
        AGh[i] = h ? ag : 0.f;
        AGv[i] = h ? 0.f : ag;
        if (in_tile(u, v)) RP[(v - 4) * TW + (u - 4)] = (gd[0] + gd[2]) + 0.5f * S;   // the own sample feeds both r and b
    }
    __syncthreads();
    // pass 4: raw adjoint of the tile's samples inside the area
    for (int i = threadIdx.x; i < TW * TH; i += kNT) {
        const int v = 4 + i / TW, u = 4 + (i - (i / TW) * TW);
        if (!in_a(u, v)) continue;
        const bool green = (u & 1) != F::nx(v & 1);
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
    flush_params(a, acc, reinterpret_cast<double*>(lds + L::OFF_RAW));
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
        if (interior) menon_back_interior<TW, TH, 1>(a, s, lds, tx, ty);
        else menon_back_border<TW, TH, 1>(a, s, lds, tx, ty);
    } else {
        if (interior) menon_back_interior<TW, TH, 0>(a, s, lds, tx, ty);
        else menon_back_border<TW, TH, 0>(a, s, lds, tx, ty);
    }
}

// ------------------------------------------- Menon, recorded decisions --
//
// Interior tiles whose adjoint reach (tile +-4) has decisions recorded by
// the forward (rawops_render_record): the classifier is not recomputed.  The
// forward values the adjoint needs besides the decisions -- the selected
// green at non-green sites of tile +-2 and red/blue at green sites of tile
// +-1, for the demosaic output of the tile's pixels (parameter gradients) --
// come from the sample codes of tile +-4, with the forward's FP64 operation
// order.  The gather passes are those of menon_back_interior.
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
    // After pass 1: AR/AB over green + red; after pass 2: AGh/AGv over PO;
    // the parameter reduction over the codes.
    static_assert(2 * NC * 4 <= 2 * GW * GH * 8, "AR/AB do not fit over green/red");
    static_assert((kNT / 64) * 9 * 8 <= OFF_HZ, "the parameter reduction does not fit over the codes");
};

template <int TW, int TH, int N0>
__device__ __forceinline__ void menon_back_recorded_tile(const BackwardArgs& a, const SensorArgs& s, char* lds,
                                                         int tx, int ty) {
    using R = RecLds<TW, TH>;
    constexpr int EW = R::EW, EH = R::EH, HWE = R::HWE, NC = R::NC, GW = R::GW;
    constexpr int NI = HWE * (EH / 2), J = (NI + kNT - 1) / kNT;   // row-pair items over the extended tile
    constexpr int NG = GW * ((EH - 4) / 2);                         // row-pair items over tile +-2
    const int ex = tx - 4, ey = ty - 4;
    uint16_t* raw = reinterpret_cast<uint16_t*>(lds + R::OFF_RAW);
    uint8_t* hz = reinterpret_cast<uint8_t*>(lds + R::OFF_HZ);
    double* green = reinterpret_cast<double*>(lds + R::OFF_GREEN);
    double* red = reinterpret_cast<double*>(lds + R::OFF_RED);
    double* blue = reinterpret_cast<double*>(lds + R::OFF_BLUE);
    auto nx = [](int b) { return (b & 1) ? 1 - N0 : N0; };          // x parity of non-green sites, row parity b
    SiteLevels lv;
    lv.init(s, ex, ey);
    uint32_t red_rows = 0;
    for (int b = 0; b < 2; ++b) red_rows |= (color_of(s, ex + nx(b), ey + b) == 0 ? 1u : 0u) << b;
    auto Rn = [&](int u, int v) -> double {    // parity of (u, v) relative to the extended tile
        return (double)lv.norm(raw[v * EW + u], v & 1, u & 1);
    };
    auto g2 = [](int u, int v) { return (v - 2) * GW + ((u - 2) >> 1); };   // tile +-2 compaction

    // codes of the extended tile, its recorded decisions, upstream gradients
    {
        const int64_t stride = s.stride;
        const uint16_t* base = s.samples + ((int64_t)ey * stride + ex);
        const bool pairs = ((reinterpret_cast<uintptr_t>(base) | (uintptr_t)(stride * 2)) & 3) == 0;
        for (int i = threadIdx.x; i < EW * EH / 2; i += kNT) {
            const int v = i / (EW / 2), k = i - v * (EW / 2);
            const uint16_t* p = base + (int64_t)v * stride + 2 * k;
            if (pairs) {
                *reinterpret_cast<uint32_t*>(raw + v * EW + 2 * k) = *reinterpret_cast<const uint32_t*>(p);
            } else {
                raw[v * EW + 2 * k] = p[0];
                raw[v * EW + 2 * k + 1] = p[1];
            }
        }
        for (int i = threadIdx.x; i < NC; i += kNT) {
            const int v = i / HWE, u = 2 * (i - v * HWE) + nx(v & 1);
            hz[i] = a.decisions[(size_t)(ey + v - s.ay) * a.dpitch + (size_t)((ex + u - s.ax) >> 1)];
        }
    }
    float goN[J][2][3], goG[J][2][3];
    uint32_t hasN = 0, hasG = 0;    // bit 2j + b
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it < NI) {
            const int qr = it / HWE, k = it - qr * HWE;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int v = 2 * qr + b, un = 2 * k + nx(b);
                if (load_gout(a, ex + un, ey + v, goN[j][b])) hasN |= 1u << (2 * j + b);
                if (load_gout(a, ex + (un ^ 1), ey + v, goG[j][b])) hasG |= 1u << (2 * j + b);
            }
        }
    }
    __syncthreads();
    // selected green at non-green sites of tile +-2 (MenonDemosaic.cpp:107-119, 141-147)
    for (int it = threadIdx.x; it < NG; it += kNT) {
        const int qr = it / GW, k = it - qr * GW;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 + 2 * qr + b, u = 2 + 2 * k + nx(b);
            const double own = Rn(u, v);
            double gsel;
            if (hz[v * HWE + (u >> 1)])
                gsel = 0.5 * (Rn(u - 1, v) + Rn(u + 1, v)) + 0.5 * own - 0.25 * (Rn(u - 2, v) + Rn(u + 2, v));
            else
                gsel = 0.5 * (Rn(u, v - 1) + Rn(u, v + 1)) + 0.5 * own - 0.25 * (Rn(u, v - 2) + Rn(u, v + 2));
            green[g2(u, v)] = gsel;
        }
    }
    __syncthreads();
    // red and blue at green sites of tile +-1 (MenonDemosaic.cpp:148-161)
    for (int it = threadIdx.x; it < NG; it += kNT) {
        const int qr = it / GW, k = it - qr * GW;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 + 2 * qr + b, u = 2 + 2 * k + (nx(b) ^ 1);
            if (u < 3 || u >= EW - 3 || v < 3 || v >= EH - 3) continue;
            const double own = Rn(u, v);
            const double ch = own + 0.5 * (Rn(u - 1, v) + Rn(u + 1, v)) - 0.5 * (green[g2(u - 1, v)] + green[g2(u + 1, v)]);
            const double cv = own + 0.5 * (Rn(u, v - 1) + Rn(u, v + 1)) - 0.5 * (green[g2(u, v - 1)] + green[g2(u, v + 1)]);
            const bool red_row = (red_rows >> b) & 1u;
            red[g2(u, v)] = red_row ? ch : cv;
            blue[g2(u, v)] = red_row ? cv : ch;
        }
    }
    __syncthreads();

    float* POh = reinterpret_cast<float*>(lds + R::OFF_PO);
    float* POv = POh + NC;
    float* AR = reinterpret_cast<float*>(lds + R::OFF_GREEN);
    float* AB = AR + NC;
    float* AGh = POh;
    float* AGv = POv;
    const bool red0 = (red_rows & 1u) != 0, red1 = (red_rows & 2u) != 0;
    auto red_row = [&](int b) { return b ? red1 : red0; };
    auto tile_item = [](int qr, int k) { return qr >= 2 && qr < 2 + TH / 2 && k >= 2 && k < 2 + TW / 2; };
    float pg[J][2], rn[J][2], gg[J][2][3], rpg[J][2], rpn[J][2];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

    // pass 1
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int it = (int)threadIdx.x + j * kNT;
        if (it >= NI) break;
        const int qr = it / HWE, k = it - qr * HWE;
        const bool mine = tile_item(qr, k);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int v = 2 * qr + b, nb = nx(b), un = 2 * k + nb, ug = un ^ 1, c = v * HWE + k;
            const bool rs = red_row(b);
            const bool hn = (hasN >> (2 * j + b)) & 1u, hg = (hasG >> (2 * j + b)) & 1u;
            float gd[3], uu[3];
            if (hn) chain_back(a, goN[j][b], uu, gd);
            else gd[0] = gd[1] = gd[2] = 0.f;
            const bool h = hz[c] != 0;
            const float po = rs ? gd[2] : gd[0];
            POh[c] = h ? po : 0.f;
            POv[c] = h ? 0.f : po;
            pg[j][b] = gd[1];
            rn[j][b] = gd[0] + gd[2];           // the own sample feeds both r and b
            if (hg) chain_back(a, goG[j][b], uu, gg[j][b]);
            else gg[j][b][0] = gg[j][b][1] = gg[j][b][2] = 0.f;
            if (mine) {
                if (hn) {
                    // final chroma (MenonDemosaic.cpp:162-173), as the forward takes it
                    const int ia = g2(un - (h ? 1 : 0), v - (h ? 0 : 1)), ic = g2(un + (h ? 1 : 0), v + (h ? 0 : 1));
                    const double mr = 0.5 * (red[ia] + red[ic]);
                    const double mb = 0.5 * (blue[ia] + blue[ic]);
                    const float self = lv.norm(raw[v * EW + un], b, nb);
                    const float other = (float)((double)self + (rs ? mb : mr) - (rs ? mr : mb));
                    const float d[3] = {rs ? self : other, (float)green[g2(un, v)], rs ? other : self};
                    param_grads(d, goN[j][b], acc);
                }
                if (hg) {
                    const int g = g2(ug, v);
                    const float d[3] = {(float)red[g], lv.norm(raw[v * EW + ug], b, nb ^ 1), (float)blue[g]};
                    param_grads(d, goG[j][b], acc);
                }
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
            const int v = 2 * qr + b, nb = nx(b), ug = 2 * k + 1 - nb, c = v * HWE + k;
            if (ug < 1 || ug >= EW - 1 || v < 1 || v >= EH - 1) continue;
            const float sh = 0.5f * (POh[c - nb] + POh[c + 1 - nb]);
            const float sv = 0.5f * (POv[c - HWE] + POv[c + HWE]);
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
            const int v = 2 * qr + b, nb = nx(b), un = 2 * k + nb, c = v * HWE + k;
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
            const int v = 2 * qr + b, nb = nx(b), un = 2 * k + nb, ug = un ^ 1, c = v * HWE + k;
            const float rgv = rpg[j][b] + 0.5f * (AGh[c - nb] + AGh[c + 1 - nb]) + 0.5f * (AGv[c - HWE] + AGv[c + HWE]);
            const float rnv = rpn[j][b] + 0.5f * (AGh[c] + AGv[c]) - 0.25f * (AGh[c - 1] + AGh[c + 1]) -
                              0.25f * (AGv[c - 2 * HWE] + AGv[c + 2 * HWE]);
            if (a.graw) {
                float* row = a.graw + (size_t)(ty + v - 4 - s.ay) * s.aw + (size_t)(tx - 4 - s.ax);
                row[un] = rnv;
                row[ug] = rgv;
            }
        }
    }
    flush_params(a, acc, reinterpret_cast<double*>(lds + R::OFF_RAW));
}

// One workgroup per tile of the recorded block [r0, r1) x [c0, c1).
template <int TW, int TH>
__global__ void __launch_bounds__(kNT) menon_backward_recorded(const BackwardArgs args) {
    __shared__ __attribute__((aligned(16))) char lds[RecLds<TW, TH>::BYTES];
    (void)args;
    const BackwardArgs& a = *(const BackwardArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    const int cols = a.c1 - a.c0;
    const int r = a.r0 + (int)blockIdx.x / cols, c = a.c0 + (int)blockIdx.x % cols;
    const int tx = __builtin_amdgcn_readfirstlane(a.s.ax + (a.tile0_x + c) * TW);
    const int ty = __builtin_amdgcn_readfirstlane(a.s.ay + (a.tile0_y + r) * TH);
    const SensorArgs& s = a.s;
    const uint32_t n0 = ((s.ng_parity >> ((uint32_t)ty & 1u)) & 1u) ^ ((uint32_t)tx & 1u);
    if (n0) menon_back_recorded_tile<TW, TH, 1>(a, s, lds, tx, ty);
    else menon_back_recorded_tile<TW, TH, 0>(a, s, lds, tx, ty);
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
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int EW = TW + 2, EH = TH + 2;
    for (int i = threadIdx.x; i < EW * EH; i += kNT) {
        const int v = i / EW, u = i - v * EW;       // tile coordinates + 1
        const int sx = tx + u - 1, sy = ty + v - 1;
        if (!in_area(s, sx, sy)) continue;
        float go[3];
        if (!load_gout(a, sx, sy, go)) continue;
        float uu[3], gd[3];
        chain_back(a, go, uu, gd);
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
    a.gout = L.g_out;
    a.gstride = L.g_stride;
    a.graw = L.g_raw;
    a.gparams = L.g_params;
    // Only raw samples within 4 of the viewport receive gradient (the
    // adjoint's reach); tiles are anchored at the active-area origin.  The
    // Menon tile is 64x20 so its recompute region (tile +-4) fits three
    // workgroups per CU in LDS.
    const bool menon = L.demosaic == RAWOPS_DEMOSAIC_MENON;
    const int TW = 64, TH = menon ? 20 : 32;
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
        // Tiles of the interior block whose reach (tile +-4) lies inside the
        // recorded viewport read the recorded decisions; the rest recompute.
        auto covered = [](int first, int T, int lo, int rec_lo, int rec_len, int& i0, int& i1) {
            int n0 = i1, n1 = i0;
            for (int i = i0; i < i1; ++i) {
                const int t0 = lo + (first + i) * T;
                if (t0 - 4 >= rec_lo && t0 + T + 4 <= rec_lo + rec_len) { n0 = std::min(n0, i); n1 = std::max(n1, i + 1); }
            }
            if (n0 >= n1) n0 = n1 = i0;
            i0 = n0; i1 = n1;
        };
        covered(a.tile0_y, TH, a.s.ay, (int)L.recorded.y, (int)L.recorded.height, a.r0, a.r1);
        covered(a.tile0_x, TW, a.s.ax, (int)L.recorded.x, (int)L.recorded.width, a.c0, a.c1);
        if (a.r0 == a.r1 || a.c0 == a.c1) a.r0 = a.r1 = a.c0 = a.c1 = 0;
        a.decisions = L.decisions;
        a.dpitch = decision_pitch((uint32_t)a.s.aw);
        const unsigned rec = (unsigned)((a.r1 - a.r0) * (a.c1 - a.c0));
        if (rec) {
            hipLaunchKernelGGL((menon_backward_recorded<64, 20>), dim3(rec), dim3(kNT), 0, stream, a);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        // the other tiles come first in the border-first order: launch only those
        if (blocks - rec)
            hipLaunchKernelGGL((menon_backward<64, 20>), dim3(blocks - rec), dim3(kNT), 0, stream, a);
        return hipGetLastError();
    }
    if (a.r0 == a.r1 || a.c0 == a.c1) a.r0 = a.r1 = a.c0 = a.c1 = 0;
    if (menon)
        hipLaunchKernelGGL((menon_backward<64, 20>), dim3(blocks), dim3(kNT), 0, stream, a);
    else
        hipLaunchKernelGGL((bilinear_backward<64, 32>), dim3(blocks), dim3(kNT), 0, stream, a);
    return hipGetLastError();
}

}  // namespace rawops
