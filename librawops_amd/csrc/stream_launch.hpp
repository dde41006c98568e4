// Launch descriptor of the streaming Menon kernel (menon_stream.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace rawops {

struct StreamArgs {
    const uint16_t* samples;
    uint32_t stride;             // samples per row
    int ax, ay, aw, ah;          // active area (loads are clamped into it)
    int X0, Y0, X1, Y1;          // rectangle rendered (sensor coordinates), 6-pixel support inside the area
    int vx, vy;                  // sensor position of output pixel (0, 0)
    int band_rows;               // rows per wavefront (a multiple of 4; 0: the default, 16)
    int wave_cols, bands;        // set by the launcher
    // Normalisation levels by row offset parity from Y0 and lane column
    // (0: the column of parity X0, 1: the other).
    float blk[2][2], rcp[2][2], den[2][2];
    const OpList* native;
    const OpList* reduced;
    uint32_t tap;                // reduced ops before the FP32 output
    float* out;                  // viewport RGB, may be null
    uint8_t* out8;               // may be null
    uint32_t ostride;            // output row stride in pixels
    uint32_t* flags;             // may be null
    // Linear chains: [white-balance gains, exposure multiplier] [3x3], in
    // the reference's node order (WhiteBalanceNode, ExposureNode,
    // CameraToWorkingNode).  A chain with only one of the first two runs the
    // other as exact factors of 1.0f (x * 1.0f == x for every float).
    float gains[3], mult, m[9];
    // Menon decisions of the viewport (the 64 x 32-tile record of
    // render_launch.hpp, zeroed by the launcher), or null.
    uint32_t* decisions;
    uint32_t dtiles;             // record tiles per viewport row
    // Approximate form: the Menon stages in FP32 instead of the reference's
    // FP64 (linear chains only; its outputs and decisions differ from the
    // reference's, the flip rate is measured, not bounded).
    int fp32;
};

// Op program forms of the streaming kernel: the device op lists run by the
// interpreter once per row pair, or a linear chain with its parameters in
// the kernel arguments (STREAM_OPS_LINEAR | the stages present).
enum StreamOps : int {
    STREAM_OPS_LISTS = 0,
    STREAM_OPS_LINEAR = 8,
    STREAM_LIN_GAINS = 1,       // gains and multiplier
    STREAM_LIN_MATRIX = 2       // 3x3
};

// n0: lane column of the non-green sites in row Y0; red0: they are red.
hipError_t launch_menon_stream(const StreamArgs& args, int n0, int red0, int ops, hipStream_t stream);

}  // namespace rawops
