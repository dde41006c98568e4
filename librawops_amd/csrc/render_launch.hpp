// Host-side launch descriptors shared between the C ABI (capi.hip) and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

#include "../../include/rawops_hip.h"
#include "device_common.hpp"

namespace rawops {

struct RenderLaunch {
    SensorArgs sensor;
    uint32_t demosaic;
    uint32_t mip;
    rawops_rect viewport;    // level coordinates
    float* out;
    uint8_t* out8;
    uint32_t ostride;
    uint32_t* flags;
    uint32_t* decisions;     // Menon mip 0: record classifier decisions (may be null)
    const OpList* native_host;   // host copies of the plan's op lists (may be null)
    const OpList* reduced_host;
};

// native / reduced / post point at device-resident OpLists (owned by a plan).

// Menon decision record of a viewport (vx, vy, vw, vh): one bit per non-green
// photosite, 1 = horizontal, in blocks of 32 words per 64 x 32 tile of the
// viewport (a tile's block is one 128-byte line, written whole).  With
// X = x - vx, Y = y - vy the site at (x, y) is bit (X >> 1) & 31 of word
// ((Y >> 5) decision_tiles(vw) + (X >> 6)) 32 + (Y & 31).
inline uint32_t decision_tiles(uint32_t width) { return (width + 63) / 64; }
inline size_t decision_bytes(uint32_t width, uint32_t height) {
    return (size_t)decision_tiles(width) * ((height + 31) / 32) * 128;
}
hipError_t launch_render(const RenderLaunch& launch, const OpList* native, const OpList* reduced,
                         hipStream_t stream);
// Menon mip-0 render path and streaming band rows (rawops_set_menon_path).
extern std::atomic<uint32_t> g_menon_path, g_stream_band;

// Adjoint of out = M * (m * (g (.) demosaic(raw))) over a native viewport.
struct BackwardLaunch {
    SensorArgs sensor;
    uint32_t demosaic;
    rawops_rect viewport;
    float gains[3], multiplier, matrix[9];
    const float* g_out;      // dL/dout, viewport RGB
    uint32_t g_stride;       // pixels per g_out row
    float* g_raw;            // dL/d(normalised raw), active-area plane (written in full)
    double* g_params;        // 13 accumulated: gains[3], multiplier, matrix[9]
    const uint32_t* decisions;  // Menon decisions recorded by a forward over `recorded` (may be null)
    rawops_rect recorded;
};

hipError_t launch_render_backward(const BackwardLaunch& launch, hipStream_t stream);

struct GuidedLaunch {
    const float* src;        // raster pixels (native resolution)
    uint32_t src_stride;     // pixels per row
    uint32_t width, height;  // native raster size
    uint32_t mip;
    rawops_rect viewport;    // level coordinates
    uint32_t radius;
    double epsilon;
    double wr, wb;           // working-Y weights
    float* out;
    uint32_t ostride;
    uint32_t* flags;
    void* workspace;         // guided_workspace_bytes() of device memory (r > 0)
    uint32_t fused_ops;      // leading post-filter ops run inside the filter pass; the rest
    uint32_t post_ops;       // (from the first curve op on, up to post_ops) run in a pass of their own
    uint32_t slices;         // a stack of this many images, each filtered on its own (0 or 1: one)
    size_t src_zstride;      // floats between consecutive images of the source stack
    size_t out_zstride;      // floats between consecutive output images
};

size_t guided_workspace_bytes(const GuidedLaunch& launch);

// Reverse mode of the working-Y guided filter at level 0 over the whole
// raster (radius >= 1).
struct GuidedBackwardLaunch {
    const float* src;        // raster p
    uint32_t src_stride, width, height;
    uint32_t radius;
    double epsilon, wr, wb;
    const float* g_out;      // dL/dq over viewport
    uint32_t g_stride;
    rawops_rect viewport;
    float* g_in;             // dL/dp, width x height x 3
    double* g_eps;           // accumulated
    void* workspace;         // guided_backward_workspace_bytes() per image of the stack
    uint32_t slices;         // a stack of this many images (0 or 1: one); dL/dp images consecutive
    size_t src_zstride;      // floats between consecutive source images
    size_t g_out_zstride;    // floats between consecutive dL/dq images
};
size_t guided_backward_workspace_bytes(uint32_t width, uint32_t height);
hipError_t launch_guided_backward(const GuidedBackwardLaunch& launch, hipStream_t stream);
hipError_t launch_guided(const GuidedLaunch& launch, const OpList* post, hipStream_t stream);

// Per-channel histogram of an FP32 RGB image, accumulated into device
// counters: counts [3][bins], tallies [3][3] (underflow, overflow,
// nonfinite), extrema [3][2] (order keys of the finite minimum and maximum).
struct HistogramLaunch {
    const float* rgb;
    uint32_t width, height, stride;   // stride in pixels
    uint32_t bins;
    double lower, upper, span;        // span = upper - lower
    unsigned long long* counts;
    unsigned long long* tallies;
    uint32_t* extrema;
    uint32_t* partial;                // set by the launcher: per-workgroup partial results, or null
    uint32_t pstride;                 // words per workgroup in `partial`
};
hipError_t launch_histogram(const HistogramLaunch& launch, hipStream_t stream);

// Local mean / population variance / finite count per channel over
// (2r+1)^2 windows clipped to the image (local_statistics_rgb,
// SpatialOps.cpp:79-121).  `in` covers `needed` (output grown by r, clipped to
// the image; row stride in pixels), outputs are output.width x output.height
// x 3, row-major.
struct LocalStatsLaunch {
    const float* in;
    uint32_t in_stride;
    rawops_rect needed, output, image;
    uint32_t radius;
    double* mean;
    double* variance;
    uint32_t* count;
};
hipError_t launch_local_stats(const LocalStatsLaunch& launch, hipStream_t stream);

hipError_t launch_srgb(const float* x, float* y, size_t n, hipStream_t s);
hipError_t launch_srgb_backward(const float* x, const float* gy, float* gx, size_t n, hipStream_t s);
hipError_t launch_exposure_backward(const float* x, const float* gy, float m, float* gx, double* g_stops, size_t n,
                                    hipStream_t s);
hipError_t launch_boost(const float* x, const float* base, float strength, float* y, size_t n, hipStream_t s);
hipError_t launch_boost_backward(const float* x, const float* base, const float* gy, float strength, float* gx,
                                 float* gbase, double* gs, size_t n, hipStream_t s);
// With g_stops set, x is the input of an exposure y = m x that feeds the
// curve: the exposure's adjoint is taken in the same pass.
hipError_t launch_curve_backward(const float* x, const float* gy, size_t n, const DevCurve* curve, uint32_t knots,
                                 uint32_t cubic, float* gx, double* gys, hipStream_t s, float m = 1.0f,
                                 double* g_stops = nullptr);

}  // namespace rawops
