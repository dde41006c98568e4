// Per-channel histogram of an FP32 RGB image (histogram_rgb,
// ImageAnalysis.cpp:10-70): counts over `bins` equal bins of [lower, upper]
// (half-open, the last bin closed), underflow / overflow / nonfinite tallies
// and the finite extrema, accumulated into device counters so that an image
// can be analysed band by band.
//
// One pass over the pixels (a band of rows per workgroup).  With up
// to kLdsBins bins a workgroup counts into LDS and adds its nonzero bins to
// the global counters once; wider histograms count with global atomics.  The
// tallies and extrema are reduced per wave with shuffles and added by one
// lane per wave.  Extrema are kept as order keys
// (unsigned integers that sort as the floats do) so that atomicMin/Max apply.
#include "render_launch.hpp"

#include <algorithm>

namespace rawops {
namespace {

constexpr int kHistThreads = 256;
constexpr uint32_t kLdsBins = 2048;

__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(kHistThreads) histogram_rgb(const HistogramLaunch h) {
    __shared__ uint32_t lh[3 * kLdsBins];
    const bool in_lds = h.bins <= kLdsBins;
    const uint32_t nb = 3 * h.bins;
    if (in_lds) {
        for (uint32_t i = threadIdx.x; i < nb; i += kHistThreads) lh[i] = 0;
        __syncthreads();
    }
    uint32_t under[3] = {0, 0, 0}, over[3] = {0, 0, 0}, nonfinite[3] = {0, 0, 0};
    uint32_t kmin[3] = {~0u, ~0u, ~0u}, kmax[3] = {0, 0, 0};
    const double bins = (double)h.bins;
    // a band of whole rows per workgroup, threads across each row
    const uint32_t rows = (h.height + gridDim.x - 1) / gridDim.x;
    const uint32_t y0 = blockIdx.x * rows, y1 = min(h.height, y0 + rows);
    for (uint32_t y = y0; y < y1; ++y)
    for (uint32_t x = threadIdx.x; x < h.width; x += kHistThreads) {
        const float* q = h.rgb + ((size_t)y * h.stride + x) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = q[c];
            if (!__builtin_isfinite(s)) { ++nonfinite[c]; continue; }
            const uint32_t k = order_key(s);
            kmin[c] = min(kmin[c], k);
            kmax[c] = max(kmax[c], k);
            if ((double)s < h.lower) { ++under[c]; continue; }
            if ((double)s > h.upper) { ++over[c]; continue; }
            // Divide before multiplying (no overflow for extreme ranges).
            const uint32_t bin = min(h.bins - 1, (uint32_t)(((double)s - h.lower) / h.span * bins));
            if (in_lds) atomicAdd(&lh[c * h.bins + bin], 1u);
            else atomicAdd(h.counts + (size_t)c * h.bins + bin, 1ull);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int off = 32; off > 0; off >>= 1) {
            under[c] += __shfl_xor(under[c], off, 64);
            over[c] += __shfl_xor(over[c], off, 64);
            nonfinite[c] += __shfl_xor(nonfinite[c], off, 64);
            kmin[c] = min(kmin[c], (uint32_t)__shfl_xor((int)kmin[c], off, 64));
            kmax[c] = max(kmax[c], (uint32_t)__shfl_xor((int)kmax[c], off, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (under[c]) atomicAdd(h.tallies + 3 * c, (unsigned long long)under[c]);
            if (over[c]) atomicAdd(h.tallies + 3 * c + 1, (unsigned long long)over[c]);
            if (nonfinite[c]) atomicAdd(h.tallies + 3 * c + 2, (unsigned long long)nonfinite[c]);
            if (kmin[c] != ~0u) atomicMin(h.extrema + 2 * c, kmin[c]);
            if (kmax[c] != 0u) atomicMax(h.extrema + 2 * c + 1, kmax[c]);
        }
    }
    if (in_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += kHistThreads)
            if (lh[i]) atomicAdd(h.counts + i, (unsigned long long)lh[i]);
    }
}

}  // namespace

hipError_t launch_histogram(const HistogramLaunch& h, hipStream_t stream) {
    const size_t n = (size_t)h.width * h.height;
    if (!n) return hipSuccess;
    // At least eight pixels per thread, and at most 512 workgroups: each adds
    // its bins to the same global counters (1024 or 2048 workgroups, or
    // 512- and 1024-thread ones, measured slower on a 24 MP frame).
    const size_t want = (n + (size_t)kHistThreads * 8 - 1) / ((size_t)kHistThreads * 8);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>({want, 512, (size_t)h.height}));
    hipLaunchKernelGGL(histogram_rgb, dim3(blocks), dim3(kHistThreads), 0, stream, h);
    return hipGetLastError();
}

}  // namespace rawops
