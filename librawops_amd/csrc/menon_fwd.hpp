// Forward Menon (DDFAPD) tile for the render kernels.
//
// A workgroup of NT threads demosaics one TW x TH tile held in LDS.  Work is
// handed out as row-pair items: an item covers the same compacted column in
// two adjacent rows, so it holds exactly one site of each row parity and the
// position of every neighbour it reads is a compile-time offset from the
// item's base address (N0, the tile-relative x parity of the non-green sites
// in even rows, is a template parameter: it is the same for every tile of a
// launch).  Arithmetic is FP64 in the reference's operation order
// (MenonDemosaic.cpp:89-174) and must be compiled with -ffp-contract=off.
//
// LDS planes (checkerboard-compacted: one value per two columns):
//   raw    u16   tile +-6, all sites: the sample codes, normalised where read
//                (the per-site levels are per-tile constants by CFA parity,
//                and the normalisation is exact, so every read returns the
//                same float; 2 bytes a sample keep three workgroups per CU)
//   gh     f64   tile +-4, non-green sites; overwritten in place by the
//                selected green (classifier output)
//   gv     f64   tile +-4, non-green sites; red at green sites afterwards,
//                or (SAMPLES) the normalised samples of the non-green sites
//                of tile +-2, stored by the classifier, with red over dv
//   dh     f64   x [-4, TW+2), y [-4, TH+4); blue at green sites afterwards
//   dv     f64   x [-4, TW+4), y [-4, TH+2)
//   hz     u8    classifier decision at the tile's non-green sites
// Tiles whose 6-pixel support lies inside the active area take the INTERIOR
// form (BORDER = false); the others reflect reads at the area edge and clip
// computation to the area exactly as the reference's clipped planes do.
#pragma once

#include "device_common.hpp"

namespace rawops {

template <int TW, int TH, bool SAMPLES = false>
struct MenonFwdLds {
    static constexpr int RW = TW + 12, RH = TH + 12;     // raw, origin (-6, -6)
    static constexpr int NP = (TW + 8) / 2, NR = TH + 8; // gh / gv, origin (-4, -4)
    static constexpr int DHP = (TW + 6) / 2, DHR = TH + 8;
    static constexpr int DVP = (TW + 8) / 2, DVR = TH + 6;
    static constexpr int GP = (TW + 4) / 2, GR = TH + 4; // red / blue, origin (-2, -2)
    static constexpr int OFF_RAW = 0;
    static constexpr int OFF_GH = (RW * RH * 2 + 15) / 16 * 16;
    static constexpr int OFF_GV = OFF_GH + NP * NR * 8;
    static constexpr int OFF_DH = OFF_GV + NP * NR * 8;
    static constexpr int OFF_DV = OFF_DH + DHP * DHR * 8;
    static constexpr int OFF_HZ = OFF_DV + DVP * DVR * 8;
    static constexpr int BYTES = OFF_HZ + (TW / 2) * TH;
    static constexpr int OFF_RED = SAMPLES ? OFF_DV : OFF_GV, OFF_BLUE = OFF_DH;
    // Native-tile staging for mip reduction: every plane is dead by then.
    static constexpr int OFF_NATIVE = 0;
    static_assert(GP * GR <= NP * NR && GP * GR <= DHP * DHR && GP * GR <= DVP * DVR,
                  "red/blue do not fit over gv or dv / dh");
    static_assert(TW * TH * 12 <= BYTES, "native staging does not fit");
    static_assert(TW % 2 == 0 && TH % 2 == 0, "tiles must cover whole Bayer cells");
    // The directional stage reads normalised FP32 samples staged over dh/dv,
    // which the gradient stage only writes afterwards: two planes by column
    // parity (RFP per row), so that the site-per-lane reads are contiguous.
    static constexpr int OFF_RAWF = OFF_DH, RFP = RW / 2, RF_PLANE = RFP * RH;
    static_assert(OFF_DV == OFF_DH + DHP * DHR * 8 && RW * RH * 4 <= DHP * DHR * 8 + DVP * DVR * 8,
                  "normalised samples do not fit over dh/dv");
};

// Per-site normalisation levels of a tile by (row, column) parity relative to
// its origin (whole-sample reflection keeps parity).  They are uniform over
// the workgroup; readfirstlane keeps them opaque scalars, so a run-time parity
// select cannot be folded into an indexed read of the kernel arguments
// (which would copy them to scratch memory).
struct SiteLevels {
    float blk[2][2], rcp[2][2], den[2][2];

    __device__ __forceinline__ void init(const SensorArgs& s, int tx, int ty) {
        auto uni = [](float v) { return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v))); };
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const uint32_t site = site_of(s, tx + e, ty + b);
                blk[b][e] = uni(site_value(s.black, site));
                rcp[b][e] = uni(site_value(s.rcp, site));
                den[b][e] = uni(site_value(s.denom, site));
            }
    }
    // Selects rather than an indexed load (see above).
    __device__ __forceinline__ static float lvl(const float (&v)[2][2], int py, int px) {
        return py ? (px ? v[1][1] : v[1][0]) : (px ? v[0][1] : v[0][0]);
    }
    // Normalised sample (RawEngine.cpp:89-97) of a code at row / column parity (py, px).
    __device__ __forceinline__ float norm(uint32_t code, int py, int px) const {
        return normalize_exact((float)code, lvl(blk, py, px), lvl(rcp, py, px), lvl(den, py, px));
    }
};

template <int TW, int TH, int NT, int N0, bool BORDER, bool SAMPLES = false>
struct MenonFwd {
    using L = MenonFwdLds<TW, TH, SAMPLES>;
    static constexpr int RW = L::RW, NP = L::NP, DHP = L::DHP, DVP = L::DVP, GP = L::GP;

    const SensorArgs& s;
    int tx, ty;             // tile origin (sensor)
    uint32_t red_rows;      // bit b: the non-green sites of tile rows with parity b are red
    uint16_t* raw;
    float* rawf;            // normalised samples by column parity, valid from load() until gradients() writes dh/dv
    double *gh, *gv, *dh, *dv, *red, *blue;
    uint8_t* hz;
    SiteLevels lv;          // levels by (row, column) parity relative to the tile origin

    __device__ __forceinline__ MenonFwd(const SensorArgs& s_, char* lds, int tx_, int ty_) : s(s_), tx(tx_), ty(ty_) {
        raw = reinterpret_cast<uint16_t*>(lds + L::OFF_RAW);
        gh = reinterpret_cast<double*>(lds + L::OFF_GH);
        gv = reinterpret_cast<double*>(lds + L::OFF_GV);
        dh = reinterpret_cast<double*>(lds + L::OFF_DH);
        rawf = reinterpret_cast<float*>(lds + L::OFF_RAWF);
        dv = reinterpret_cast<double*>(lds + L::OFF_DV);
        red = reinterpret_cast<double*>(lds + L::OFF_RED);
        blue = reinterpret_cast<double*>(lds + L::OFF_BLUE);
        hz = reinterpret_cast<uint8_t*>(lds + L::OFF_HZ);
        red_rows = 0;
        for (int b = 0; b < 2; ++b) red_rows |= (color_of(s, tx + nx(b), ty + b) == 0 ? 1u : 0u) << b;
        lv.init(s, tx, ty);
    }

    // Normalised sample at LDS index i of a position with row parity py and
    // column parity px (relative to the tile origin).
    __device__ __forceinline__ float Rn(int i, int py, int px) const { return lv.norm(raw[i], py, px); }

    // x parity (tile-relative) of the non-green sites in rows of parity b.
    __device__ static constexpr int nx(int b) { return (b & 1) ? 1 - N0 : N0; }

    __device__ __forceinline__ bool inside(int rx, int ry) const { return !BORDER || in_area(s, tx + rx, ty + ry); }
    __device__ __forceinline__ void mirror(int& rx, int& ry) const {
        if (BORDER) {
            const int sx = tx + rx, sy = ty + ry;
            if (!in_area(s, sx, sy)) {
                rx = reflect(sx, s.ax, s.aw) - tx;
                ry = reflect(sy, s.ay, s.ah) - ty;
            }
        }
    }
    __device__ __forceinline__ static int rawi(int rx, int ry) { return (ry + 6) * RW + rx + 6; }
    __device__ __forceinline__ static int rfi(int rx, int ry) {
        return ((rx + 6) & 1) * L::RF_PLANE + (ry + 6) * L::RFP + ((rx + 6) >> 1);
    }
    __device__ __forceinline__ static int ni(int rx, int ry) { return (ry + 4) * NP + ((rx + 4) >> 1); }
    __device__ __forceinline__ static int gi(int rx, int ry) { return (ry + 2) * GP + ((rx + 2) >> 1); }

    // ------------------------------------------------------------- raw --
    // Normalised samples of tile +-6 (RawEngine.cpp:89-97 / MenonDemosaic.cpp:81-86)
    // as 2x2 cells: the four sites of a cell have fixed CFA positions, so
    // their levels are per-tile constants.
    __device__ __forceinline__ void load() {
        constexpr int QK = (TW + 12) / 2, QR = (TH + 12) / 2, N = QK * QR;
        const int64_t stride = s.stride;
        const uint16_t* base = s.samples + ((int64_t)(ty - 6) * stride + (tx - 6));
        const bool pairs = !BORDER && ((reinterpret_cast<uintptr_t>(base) | (uintptr_t)(stride * 2)) & 3) == 0;
        constexpr int J = (N + NT - 1) / NT;
        auto put = [&](int qr, int qk, int b, int e, uint32_t code) {
            const int i = (2 * qr + b) * RW + 2 * qk + e;
            raw[i] = (uint16_t)code;
            rawf[e * L::RF_PLANE + (2 * qr + b) * L::RFP + qk] = lv.norm(code, b, e);
        };
        if (pairs) {
            // Interior, 4-byte aligned cells: issue every load first, then convert.
            uint32_t w[J][2];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int it = (int)threadIdx.x + j * NT;
                const int qr = it / QK, qk = it - qr * QK;
                const uint16_t* p = base + ((int64_t)(2 * qr) * stride + 2 * qk);
                if (it < N) {
                    w[j][0] = *reinterpret_cast<const uint32_t*>(p);
                    w[j][1] = *reinterpret_cast<const uint32_t*>(p + stride);
                }
            }
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int it = (int)threadIdx.x + j * NT;
                const int qr = it / QK, qk = it - qr * QK;
                if (it < N) {
#pragma unroll
                    for (int b = 0; b < 2; ++b) {   // both codes of the cell row in one 4-byte store
                        const int i = (2 * qr + b) * RW + 2 * qk;
                        *reinterpret_cast<uint32_t*>(raw + i) = w[j][b];
                        const int f = (2 * qr + b) * L::RFP + qk;
                        rawf[f] = lv.norm(w[j][b] & 0xffffu, b, 0);
                        rawf[L::RF_PLANE + f] = lv.norm(w[j][b] >> 16, b, 1);
                    }
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            if (it >= N) break;
            const int qr = it / QK, qk = it - qr * QK;
            const uint16_t* p = base + ((int64_t)(2 * qr) * stride + 2 * qk);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const bool ok = !BORDER || in_area(s, tx - 6 + 2 * qk + e, ty - 6 + 2 * qr + b);
                    if (ok) put(qr, qk, b, e, p[b * stride + e]);
                }
        }
    }

    // ----------------------------------------------------- directional --
    // Directional greens at non-green sites, tile +-4 (MenonDemosaic.cpp:107-119).
    // The samples the gradient stage reads (own, x + 2, y + 2) can be kept
    // in registers across the barrier (Carry): the staged floats lie over
    // dh / dv, which that stage writes.  Both stages walk the same items.
    static constexpr int JD = (NP * ((TH + 8) / 2) + NT - 1) / NT;
    struct Carry { float own[JD][2], h2[JD][2], v2[JD][2]; };

    __device__ __forceinline__ void directional() { Carry c; directional(c); }
    __device__ __forceinline__ void directional(Carry& c) {
        constexpr int QK = NP, QR = (TH + 8) / 2, N = QK * QR;
#pragma unroll
        for (int j = 0; j < JD; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            if (it >= N) break;
            const int qr = it / QK, k = it - qr * QK;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int ry = -4 + 2 * qr + b, rx = -4 + 2 * k + nx(b);
                if (!inside(rx, ry)) continue;
                const int f0 = (ry + 6) * L::RFP + k;   // rx + 6 = 2 k + 2 + nx(b)
                auto F = [&](int ox, int oy) -> float {
                    if constexpr (BORDER) {
                        int x = rx + ox, y = ry + oy;
                        mirror(x, y);
                        return rawf[rfi(x, y)];
                    } else {
                        const int nb = nx(b);
                        return rawf[((nb + ox) & 1) * L::RF_PLANE + f0 + oy * L::RFP + ((nb + 2 + ox) >> 1)];
                    }
                };
                auto R = [&](int ox, int oy) -> double { return (double)F(ox, oy); };
                const float fo = F(0, 0), fh = F(2, 0), fv = F(0, 2);
                c.own[j][b] = fo;
                c.h2[j][b] = fh;
                c.v2[j][b] = fv;
                const double own = fo;
                const double h = 0.5 * (R(-1, 0) + R(1, 0)) + 0.5 * own - 0.25 * (R(-2, 0) + (double)fh);
                const double v = 0.5 * (R(0, -1) + R(0, 1)) + 0.5 * own - 0.25 * (R(0, -2) + (double)fv);
                const int n0 = (2 * qr + b) * NP + k;
                gh[n0] = h;
                gv[n0] = v;
            }
        }
    }

    // ------------------------------------------------------- gradients --
    // dh = |ch - ch(x+2)|, dv = |cv - cv(y+2)| with c = own - directional
    // green (MenonDemosaic.cpp:117-125); samples from the directional
    // stage's Carry, or (without one) normalised again from the codes.
    __device__ __forceinline__ void gradients() { gradients_impl<false>(nullptr); }
    __device__ __forceinline__ void gradients(const Carry& c) { gradients_impl<true>(&c); }

    template <bool CARRIED>
    __device__ __forceinline__ void gradients_impl(const Carry* c) {
        constexpr int QK = NP, QR = (TH + 8) / 2, N = QK * QR;
#pragma unroll
        for (int j = 0; j < JD; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            if (it >= N) break;
            const int qr = it / QK, k = it - qr * QK;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int ry = -4 + 2 * qr + b, rx = -4 + 2 * k + nx(b);
                const bool do_h = k < DHP, do_v = 2 * qr + b < TH + 6;
                if (!inside(rx, ry)) continue;
                const int r0 = rawi(rx, ry), n0 = (2 * qr + b) * NP + k, nb = nx(b);
                const double own = CARRIED ? (double)c->own[j][b] : (double)Rn(r0, b, nb);
                if (do_h) {
                    double there;
                    if constexpr (BORDER) {
                        int x = rx + 2, y = ry;
                        mirror(x, y);
                        there = (CARRIED ? (double)c->h2[j][b] : (double)Rn(rawi(x, y), b, nb)) - gh[ni(x, y)];
                    } else {
                        there = (CARRIED ? (double)c->h2[j][b] : (double)Rn(r0 + 2, b, nb)) - gh[n0 + 1];
                    }
                    dh[(2 * qr + b) * DHP + k] = fabs((own - gh[n0]) - there);
                }
                if (do_v) {
                    double there;
                    if constexpr (BORDER) {
                        int x = rx, y = ry + 2;
                        mirror(x, y);
                        there = (CARRIED ? (double)c->v2[j][b] : (double)Rn(rawi(x, y), b, nb)) - gv[ni(x, y)];
                    } else {
                        there = (CARRIED ? (double)c->v2[j][b] : (double)Rn(r0 + 2 * RW, b, nb)) - gv[n0 + 2 * NP];
                    }
                    dv[(2 * qr + b) * DVP + k] = fabs((own - gv[n0]) - there);
                }
            }
        }
    }

    // -------------------------------------------------------- classify --
    // Posterior classifier at non-green sites, tile +-2 (MenonDemosaic.cpp:126-147),
    // terms in the reference's order.  Gradients are >= +0, so the first
    // term equals 0 + term and weight-1 terms are added unscaled.  The
    // selected green overwrites gh in place (only this site reads it).
    __device__ __forceinline__ void classify() {
        constexpr int QK = (TW + 4) / 2, QR = (TH + 4) / 2, N = QK * QR;
#pragma unroll
        for (int j = 0; j < (N + NT - 1) / NT; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            if (it >= N) break;
            const int qr = it / QK, k = it - qr * QK;
            // Both sites' sums before any store: the two windows share four
            // of their eight dh and dv reads.
            bool hor[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int nb = nx(b);
                const int ry = -2 + 2 * qr + b, rx = -2 + 2 * k + nb;
                hor[b] = false;
                if (!inside(rx, ry)) continue;
                const int h0 = (2 * qr + 2 + b) * DHP + k, v0 = (2 * qr + 2 + b) * DVP + k;
                auto DH = [&](int ox, int oy) -> double {
                    if constexpr (BORDER) {
                        if (!inside(rx + ox, ry + oy)) return 0.0;
                        return dh[(ry + oy + 4) * DHP + ((rx + ox + 4) >> 1)];
                    } else {
                        return dh[h0 + oy * DHP + ((2 + nb + ox) >> 1)];
                    }
                };
                auto DV = [&](int ox, int oy) -> double {
                    if constexpr (BORDER) {
                        if (!inside(rx + ox, ry + oy)) return 0.0;
                        return dv[(ry + oy + 4) * DVP + ((rx + ox + 4) >> 1)];
                    } else {
                        return dv[v0 + oy * DVP + ((2 + nb + ox) >> 1)];
                    }
                };
                double h = DH(0, 2), v = DV(2, 0);
                h += DH(-2, 2);        v += DV(2, -2);
                h += DH(-1, 1);        v += DV(1, -1);
                h += 3.0 * DH(0, 0);   v += 3.0 * DV(0, 0);
                h += 3.0 * DH(-2, 0);  v += 3.0 * DV(0, -2);
                h += DH(-1, -1);       v += DV(-1, -1);
                h += DH(0, -2);        v += DV(-2, 0);
                h += DH(-2, -2);       v += DV(-2, -2);
                hor[b] = v >= h;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int nb = nx(b);
                const int ry = -2 + 2 * qr + b, rx = -2 + 2 * k + nb;
                if (!inside(rx, ry)) continue;
                const bool horizontal = hor[b];
                const int n0 = (2 * qr + 2 + b) * NP + k + 1;
                gh[n0] = horizontal ? gh[n0] : gv[n0];
                // Only this site reads gv here: it takes the site's sample for the chroma stage.
                if constexpr (SAMPLES) gv[n0] = (double)Rn(rawi(rx, ry), b, nb);
                if (rx >= 0 && rx < TW && ry >= 0 && ry < TH) hz[ry * (TW / 2) + (rx >> 1)] = horizontal ? 1 : 0;
            }
        }
    }

    // ---------------------------------------------------------- chroma --
    // Red and blue at green sites, tile +-1 (MenonDemosaic.cpp:148-161):
    // c = green + mean(raw) - mean(green) along the row (cH) and the column
    // (cV); red takes the direction of the red neighbours.  The four
    // neighbours are non-green sites (of tile +-2): with SAMPLES their
    // samples come from gv.
    __device__ __forceinline__ void chroma() {
        constexpr int QK = (TW + 2) / 2, QR = (TH + 2) / 2, N = QK * QR;
#pragma unroll
        for (int j = 0; j < (N + NT - 1) / NT; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            if (it >= N) break;
            const int qr = it / QK, k = it - qr * QK;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int p = (b + 1) & 1;             // parity of row -1 + 2 qr + b
                const int np = nx(p);
                const int ry = -1 + 2 * qr + b, rx = -1 + 2 * k + np;
                if (!inside(rx, ry)) continue;
                const int r0 = rawi(rx, ry);
                const int g0 = (2 * qr + 3 + b) * NP + k;   // row of this site in the gh/green plane
                auto R = [&](int ox, int oy) -> double {
                    const int py = (p + oy) & 1, px = (1 - np + ox) & 1;   // rx = -1 + 2k + np
                    if (SAMPLES && (ox != 0 || oy != 0)) {
                        if constexpr (BORDER) {
                            int x = rx + ox, y = ry + oy;
                            mirror(x, y);
                            return gv[ni(x, y)];
                        } else {
                            return gv[g0 + oy * NP + ((3 + np + ox) >> 1)];
                        }
                    }
                    if constexpr (BORDER) {
                        int x = rx + ox, y = ry + oy;
                        mirror(x, y);
                        return (double)Rn(rawi(x, y), py, px);
                    } else {
                        return (double)Rn(r0 + oy * RW + ox, py, px);
                    }
                };
                auto GR = [&](int ox, int oy) -> double {
                    if constexpr (BORDER) {
                        int x = rx + ox, y = ry + oy;
                        mirror(x, y);
                        return gh[ni(x, y)];
                    } else {
                        return gh[g0 + oy * NP + ((3 + np + ox) >> 1)];
                    }
                };
                const double own = R(0, 0);
                const double ch = own + 0.5 * (R(-1, 0) + R(1, 0)) - 0.5 * (GR(-1, 0) + GR(1, 0));
                const double cv = own + 0.5 * (R(0, -1) + R(0, 1)) - 0.5 * (GR(0, -1) + GR(0, 1));
                const bool red_row = (red_rows >> p) & 1u;
                const int o = (2 * qr + 1 + b) * GP + k + np;
                red[o] = red_row ? ch : cv;
                blue[o] = red_row ? cv : ch;
            }
        }
    }

    // ---------------------------------------------------------- finish --
    // Final RGB of the tile (MenonDemosaic.cpp:162-173) as 2x2 cells:
    // px[(2 j + b) 2 + e] is pixel (2 qx + e, 2 qy + b) of cell
    // threadIdx.x + j NT.  Pixels outside the active area are left at zero.
    template <int NPX>
    __device__ __forceinline__ void finish(float (&px)[NPX][3]) const {
        constexpr int NQ = NPX / 4;
        static_assert(NQ * NT == (TW / 2) * (TH / 2), "cells must divide evenly over the workgroup");
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int it = (int)threadIdx.x + j * NT;
            const int qy = it / (TW / 2), qx = it - qy * (TW / 2);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int nb = nx(b);
                const int ry = 2 * qy + b, xg = 2 * qx + 1 - nb, xn = 2 * qx + nb;
                float gr = 0.f, gg = 0.f, gb = 0.f, nr = 0.f, ng = 0.f, nbl = 0.f;
                const int gbase = (2 * qy + 2 + b) * GP + qx;   // red/blue row of this pixel row
                if (inside(xg, ry)) {
                    gr = (float)red[gbase + 1];
                    gg = Rn(rawi(xg, ry), b, 1 - nb);
                    gb = (float)blue[gbase + 1];
                }
                if (inside(xn, ry)) {
                    const bool horizontal = hz[ry * (TW / 2) + qx] != 0;
                    int ia, ic;
                    if constexpr (BORDER) {
                        const int dx = horizontal ? 1 : 0, dy = horizontal ? 0 : 1;
                        int xa = xn - dx, ya = ry - dy, xc = xn + dx, yc = ry + dy;
                        mirror(xa, ya);
                        mirror(xc, yc);
                        ia = gi(xa, ya);
                        ic = gi(xc, yc);
                    } else {
                        // Row neighbours x -+ 1 sit at columns qx + nb, qx + 1 + nb; column
                        // neighbours y -+ 1 at column qx + 1 of the rows above and below.
                        ia = gbase + (horizontal ? nb : 1 - GP);
                        ic = gbase + (horizontal ? 1 + nb : 1 + GP);
                    }
                    const double mr = 0.5 * (red[ia] + red[ic]);
                    const double mb = 0.5 * (blue[ia] + blue[ic]);
                    const float self = SAMPLES ? (float)gv[(2 * qy + 4 + b) * NP + qx + 2] : Rn(rawi(xn, ry), b, nb);
                    const double own = (double)self;
                    const bool red_site = (red_rows >> b) & 1u;
                    // R site: bb = rr + mean(blue) - mean(red); B site: rr = bb + mean(red) - mean(blue).
                    const float other = (float)(own + (red_site ? mb : mr) - (red_site ? mr : mb));
                    nr = red_site ? self : other;
                    ng = (float)gh[(2 * qy + 4 + b) * NP + qx + 2];
                    nbl = red_site ? other : self;
                }
                float* G = px[(2 * j + b) * 2 + 1 - nb];
                float* P = px[(2 * j + b) * 2 + nb];
                G[0] = gr; G[1] = gg; G[2] = gb;
                P[0] = nr; P[1] = ng; P[2] = nbl;
            }
        }
    }
};

}  // namespace rawops
