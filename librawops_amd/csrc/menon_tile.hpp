// Menon (DDFAPD) and bilinear tile stages shared by the forward render and
// the adjoint kernels.  A workgroup of NT threads stages one TW x TH tile
// (plus halo) in LDS; see render.hip for the data flow.  Menon arithmetic is
// FP64 in the reference's operation order (MenonDemosaic.cpp:89-174): this
// header must be compiled with -ffp-contract=off.
#pragma once

#include "device_common.hpp"
#include "../../include/rawops_hip.h"

namespace rawops {

// --------------------------------------------------------------- Menon --

template <int TW, int TH>
struct MenonLds {
    static constexpr int RAW_W = TW + 12, RAW_H = TH + 12;
    static constexpr int G4_HW = (TW + 8) / 2, G4_H = TH + 8;     // gh/gv: tile +-4
    static constexpr int DH_HW = (TW + 6) / 2, DH_H = TH + 8;     // dh: x [-4, TW+2), y [-4, TH+4)
    static constexpr int DV_HW = (TW + 8) / 2, DV_H = TH + 6;     // dv: x [-4, TW+4), y [-4, TH+2)
    static constexpr int G2_HW = (TW + 4) / 2, G2_H = TH + 4;     // green / red / blue: tile +-2
    static constexpr int RAW_BYTES = RAW_W * RAW_H * 4;
    static constexpr int G4_BYTES = G4_HW * G4_H * 8;
    static constexpr int DH_BYTES = DH_HW * DH_H * 8;
    static constexpr int DV_BYTES = DV_HW * DV_H * 8;
    static constexpr int G2_BYTES = G2_HW * G2_H * 8;
    static constexpr int HZ_BYTES = (TW / 2) * TH;
    static constexpr int OFF_RAW = 0;
    static constexpr int OFF_GH = OFF_RAW + RAW_BYTES;
    static constexpr int OFF_GV = OFF_GH + G4_BYTES;
    static constexpr int OFF_DH = OFF_GV + G4_BYTES;
    static constexpr int OFF_DV = OFF_DH + DH_BYTES;
    static constexpr int OFF_GREEN = OFF_DV + DV_BYTES;
    static constexpr int OFF_HZ = OFF_GREEN + G2_BYTES;
    static constexpr int BYTES = OFF_HZ + HZ_BYTES;
    // Native-tile staging for mip reduction aliases dh/dv/green (dead by then).
    static constexpr int OFF_NATIVE = OFF_DH;
    static_assert(TW * TH * 12 <= DH_BYTES + DV_BYTES + G2_BYTES, "native staging does not fit");
    static_assert(G2_HW * G2_H <= G4_HW * G4_H, "red/blue do not fit over gh/gv");
    static_assert(OFF_GH % 8 == 0, "double planes must be 8-byte aligned");
};

// Tile-relative coordinates: (rx, ry) = sensor - tile origin.
template <int TW, int TH, bool BORDER, int NT = 256>
struct MenonTile {
    using L = MenonLds<TW, TH>;
    const SensorArgs& s;
    int tx, ty;          // tile origin (sensor)
    uint32_t ngr;        // bit b: tile-relative x parity of non-green sites in rows with ry parity b
    float* raw;
    double *gh, *gv, *dh, *dv, *green, *red, *blue;
    uint8_t* hz;

    __device__ __forceinline__ int ng_parity(int ry) const { return (int)((ngr >> ((uint32_t)ry & 1u)) & 1u); }

    // Reflect a tile-relative position into the active area (BORDER only).
    __device__ __forceinline__ void mirror(int& rx, int& ry) const {
        if (BORDER) {
            const int sx = tx + rx, sy = ty + ry;
            if (!in_area(s, sx, sy)) {
                rx = reflect(sx, s.ax, s.aw) - tx;
                ry = reflect(sy, s.ay, s.ah) - ty;
            }
        }
    }
    __device__ __forceinline__ bool inside(int rx, int ry) const { return !BORDER || in_area(s, tx + rx, ty + ry); }

    __device__ __forceinline__ double R(int rx, int ry) const { return (double)raw[(ry + 6) * L::RAW_W + (rx + 6)]; }
    __device__ __forceinline__ double Rm(int rx, int ry) const { mirror(rx, ry); return R(rx, ry); }
    __device__ __forceinline__ static int i4(int rx, int ry) { return (ry + 4) * L::G4_HW + ((rx + 4) >> 1); }
    __device__ __forceinline__ static int idh(int rx, int ry) { return (ry + 4) * L::DH_HW + ((rx + 4) >> 1); }
    __device__ __forceinline__ static int idv(int rx, int ry) { return (ry + 4) * L::DV_HW + ((rx + 4) >> 1); }
    __device__ __forceinline__ static int i2(int rx, int ry) { return (ry + 2) * L::G2_HW + ((rx + 2) >> 1); }
    __device__ __forceinline__ double GREENm(int rx, int ry) const { mirror(rx, ry); return green[i2(rx, ry)]; }
    __device__ __forceinline__ double REDm(int rx, int ry) const { mirror(rx, ry); return red[i2(rx, ry)]; }
    __device__ __forceinline__ double BLUEm(int rx, int ry) const { mirror(rx, ry); return blue[i2(rx, ry)]; }
    // ch/cv at a non-green site: own - directional green (MenonDemosaic.cpp:117-118).
    __device__ __forceinline__ double CHm(int rx, int ry) const { mirror(rx, ry); return R(rx, ry) - gh[i4(rx, ry)]; }
    __device__ __forceinline__ double CVm(int rx, int ry) const { mirror(rx, ry); return R(rx, ry) - gv[i4(rx, ry)]; }
    __device__ __forceinline__ double DHz(int rx, int ry) const { return inside(rx, ry) ? dh[idh(rx, ry)] : 0.0; }
    __device__ __forceinline__ double DVz(int rx, int ry) const { return inside(rx, ry) ? dv[idv(rx, ry)] : 0.0; }

    // Directional green estimates at non-green sites, tile +-4 (MenonDemosaic.cpp:107-119).
    __device__ __forceinline__ void directional() {
        constexpr int HW = (TW + 8) / 2, N = HW * (TH + 8);
        for (int i = threadIdx.x; i < N; i += NT) {
            const int row = i / HW, k = i - row * HW;
            const int ry = row - 4;
            const int rx = -4 + 2 * k + ng_parity(ry);
            if (!inside(rx, ry)) continue;
            const double own = R(rx, ry);
            const double h = 0.5 * (Rm(rx - 1, ry) + Rm(rx + 1, ry)) + 0.5 * own
                           - 0.25 * (Rm(rx - 2, ry) + Rm(rx + 2, ry));
            const double v = 0.5 * (Rm(rx, ry - 1) + Rm(rx, ry + 1)) + 0.5 * own
                           - 0.25 * (Rm(rx, ry - 2) + Rm(rx, ry + 2));
            gh[i] = h;
            gv[i] = v;
        }
    }

    // Forward colour-difference gradients (MenonDemosaic.cpp:120-125).
    __device__ __forceinline__ void gradients() {
        {
            constexpr int HW = (TW + 6) / 2, N = HW * (TH + 8);
            for (int i = threadIdx.x; i < N; i += NT) {
                const int row = i / HW, k = i - row * HW;
                const int ry = row - 4;
                const int rx = -4 + 2 * k + ng_parity(ry);
                if (!inside(rx, ry)) continue;
                const double here = R(rx, ry) - gh[i4(rx, ry)];
                dh[i] = fabs(here - CHm(rx + 2, ry));
            }
        }
        {
            constexpr int HW = (TW + 8) / 2, N = HW * (TH + 6);
            for (int i = threadIdx.x; i < N; i += NT) {
                const int row = i / HW, k = i - row * HW;
                const int ry = row - 4;
                const int rx = -4 + 2 * k + ng_parity(ry);
                if (!inside(rx, ry)) continue;
                const double here = R(rx, ry) - gv[i4(rx, ry)];
                dv[i] = fabs(here - CVm(rx, ry + 2));
            }
        }
    }

    // Posterior classifier and green selection at non-green sites, tile +-2
    // (MenonDemosaic.cpp:126-147).  Weight-1 taps add the gradient directly
    // (1.0 * x == x exactly); weight-3 taps multiply.
    __device__ __forceinline__ void classify() {
        constexpr int HW = (TW + 4) / 2, N = HW * (TH + 4);
        for (int i = threadIdx.x; i < N; i += NT) {
            const int row = i / HW, k = i - row * HW;
            const int ry = row - 2;
            const int rx = -2 + 2 * k + ng_parity(ry);
            if (!inside(rx, ry)) continue;
            double h = 0.0, v = 0.0;
            h += DHz(rx, ry + 2);          v += DVz(rx + 2, ry);
            h += DHz(rx - 2, ry + 2);      v += DVz(rx + 2, ry - 2);
            h += DHz(rx - 1, ry + 1);      v += DVz(rx + 1, ry - 1);
            h += 3.0 * DHz(rx, ry);        v += 3.0 * DVz(rx, ry);
            h += 3.0 * DHz(rx - 2, ry);    v += 3.0 * DVz(rx, ry - 2);
            h += DHz(rx - 1, ry - 1);      v += DVz(rx - 1, ry - 1);
            h += DHz(rx, ry - 2);          v += DVz(rx - 2, ry);
            h += DHz(rx - 2, ry - 2);      v += DVz(rx - 2, ry - 2);
            const bool horizontal = v >= h;
            green[i] = horizontal ? gh[i4(rx, ry)] : gv[i4(rx, ry)];
            if (rx >= 0 && rx < TW && ry >= 0 && ry < TH) hz[ry * (TW / 2) + (rx >> 1)] = horizontal ? 1 : 0;
        }
    }

    // Red and blue at green sites, tile +-1 (MenonDemosaic.cpp:148-161).
    __device__ __forceinline__ void chroma() {
        constexpr int HW = (TW + 4) / 2, N = HW * (TH + 4);
        for (int i = threadIdx.x; i < N; i += NT) {
            const int row = i / HW, k = i - row * HW;
            const int ry = row - 2;
            const int rx = -2 + 2 * k + (ng_parity(ry) ^ 1);
            if (rx < -1 || rx > TW || ry < -1 || ry > TH || !inside(rx, ry)) continue;
            const bool red_horizontal = color_of(s, tx + rx + 1, ty + ry) == 0;
            const int dx = red_horizontal ? 1 : 0, dy = red_horizontal ? 0 : 1;
            const double own = R(rx, ry);
            const double mr1 = 0.5 * (Rm(rx - dx, ry - dy) + Rm(rx + dx, ry + dy));
            const double mg1 = 0.5 * (GREENm(rx - dx, ry - dy) + GREENm(rx + dx, ry + dy));
            const double mr2 = 0.5 * (Rm(rx - dy, ry - dx) + Rm(rx + dy, ry + dx));
            const double mg2 = 0.5 * (GREENm(rx - dy, ry - dx) + GREENm(rx + dy, ry + dx));
            red[i] = own + mr1 - mg1;
            blue[i] = own + mr2 - mg2;
        }
    }

    // Final RGB of one pixel (MenonDemosaic.cpp:162-173).
    __device__ __forceinline__ void finish(int rx, int ry, float& r, float& g, float& b) const {
        const int own = color_of(s, tx + rx, ty + ry);
        if (own == 1) {
            r = (float)red[i2(rx, ry)];
            g = raw[(ry + 6) * L::RAW_W + (rx + 6)];
            b = (float)blue[i2(rx, ry)];
            return;
        }
        const bool horizontal = hz[ry * (TW / 2) + (rx >> 1)] != 0;
        const int dx = horizontal ? 1 : 0, dy = horizontal ? 0 : 1;
        const double own_v = R(rx, ry);
        const double mb = 0.5 * (BLUEm(rx - dx, ry - dy) + BLUEm(rx + dx, ry + dy));
        const double mr = 0.5 * (REDm(rx - dx, ry - dy) + REDm(rx + dx, ry + dy));
        g = (float)green[i2(rx, ry)];
        // R site: bb = rr + mean(blue) - mean(red); B site: rr = bb + mean(red) - mean(blue).
        const bool red_site = own == 0;
        const float other = (float)(own_v + (red_site ? mb : mr) - (red_site ? mr : mb));
        const float self = (float)own_v;
        r = red_site ? self : other;
        b = red_site ? other : self;
    }
};

// Stage normalised samples of [tile - H, tile + TW + H) into LDS.  Only
// active-area samples are read: padding and inactive photosites never are.
template <int TW, int TH, int H, bool BORDER, int NT = 256>
__device__ __forceinline__ void load_raw(const SensorArgs& s, int tx, int ty, float* raw) {
    constexpr int W = TW + 2 * H, N = W * (TH + 2 * H);
    for (int i = threadIdx.x; i < N; i += NT) {
        const int row = i / W, col = i - row * W;
        const int sx = tx - H + col, sy = ty - H + row;
        if (BORDER && !in_area(s, sx, sy)) continue;
        const uint32_t code = s.samples[(size_t)sy * s.stride + (size_t)sx];
        raw[i] = normalize_code(s, code, site_of(s, sx, sy));
    }
}

// ------------------------------------------------------------- bilinear --

// Same-colour 3x3 mean clipped to the active area (RawEngine.cpp:316-339),
// accumulated in the reference's dy-major, dx-minor order.
template <int TW, int TH, bool BORDER>
__device__ __forceinline__ void bilinear_pixel(const SensorArgs& s, const float* raw, int tx, int ty, int rx, int ry,
                               float& r, float& g, float& b) {
    constexpr int W = TW + 2;
    const int sx = tx + rx, sy = ty + ry;
    const int own = color_of(s, sx, sy);
    float v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (own == c) { v[c] = raw[(ry + 1) * W + (rx + 1)]; continue; }
        float sum = 0.0f;
        int count = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                if (BORDER && !in_area(s, sx + dx, sy + dy)) continue;
                if (color_of(s, sx + dx, sy + dy) != c) continue;
                sum += raw[(ry + 1 + dy) * W + (rx + 1 + dx)];
                ++count;
            }
        v[c] = count ? sum / (float)count : 0.0f;
    }
    r = v[0]; g = v[1]; b = v[2];
}

}  // namespace rawops
