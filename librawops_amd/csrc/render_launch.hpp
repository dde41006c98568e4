// Host-side launch descriptors shared between the C ABI (capi.hip) and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

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
};

// native / reduced / post point at device-resident OpLists (owned by a plan).
hipError_t launch_render(const RenderLaunch& launch, const OpList* native, const OpList* reduced,
                         hipStream_t stream);

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
};

size_t guided_workspace_bytes(const GuidedLaunch& launch);
hipError_t launch_guided(const GuidedLaunch& launch, const OpList* post, hipStream_t stream);

}  // namespace rawops
