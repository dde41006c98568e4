// Raw-sample tile helpers shared by the render tile and the adjoint
// kernels: staging normalised samples of a tile and its halo into LDS, and
// the bilinear demosaic of one pixel from them.
#pragma once

#include "device_common.hpp"
#include "../../include/rawops_hip.h"

namespace rawops {

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
