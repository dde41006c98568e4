// Host/device description of one request of a batched launch.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace rawops {

struct BatchItem {
    SensorArgs sensor;
    const OpList* native;
    const OpList* reduced;
    uint32_t tap;            // reduced ops before the FP32 output
    int demosaic, mip;
    int vx, vy, vw, vh, ox0, oy0;
    int tiles_x, tiles, first;   // tiles of this request; first item number
    float* out;
    uint8_t* out8;
    uint32_t ostride;
    uint32_t* flags;
};

// A tile shape of the batched launch: one 512-thread workgroup per
// (request, tile) item running the one-request render kernel's tile.
struct BatchShape { int tw, th; };
constexpr int kBatchShapeCount = 3;
extern const BatchShape kBatchShapes[kBatchShapeCount];

size_t batch_state_bytes();
// Launch start time (s_memrealtime ticks) held by a host copy of the state.
unsigned long long batch_start_time(const void* state_host);
hipError_t launch_batch(int shape, const BatchItem* items, int n, int total, void* state, uint32_t* remaining,
                        unsigned long long* done, hipStream_t stream);

}  // namespace rawops
