// Batched heterogeneous render: many requests (different sources, viewports,
// mip levels, demosaicers and op programs) in one launch.
//
// The batch is flattened into a work list of (request, tile) items; each
// request owns a contiguous range of item numbers.  One 512-thread workgroup
// per item runs the one-request render kernel's fused tile (finding its
// request by binary search over the requests' first item numbers), and the
// workgroup that finishes a request's last tile stamps its completion time
// (s_memrealtime, 100 MHz).  Counter reset + tile launch are captured in a
// hipGraph and replayed per batch.  The host-side autotuner times the
// compiled tile shapes on the actual batch and keeps the fastest.
#include "render_tile.hpp"
#include "batch_launch.hpp"

namespace rawops {
namespace {

constexpr int cmax(int a, int b) { return a > b ? a : b; }

struct BatchState {
    unsigned long long t0;
};

__global__ void batch_reset(BatchState* st, const BatchItem* items, int n, uint32_t* remaining) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->t0 = wall_clock64();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) remaining[i] = items[i].tiles;
}

// LDS of a TW x TH tile for either demosaicer; the 64x32 tile fits three
// workgroups per CU (six waves per SIMD), the larger ones one.
template <int TW, int TH>
struct BatchTile {
    static constexpr int LDS = cmax(Lds<RAWOPS_DEMOSAIC_MENON, TW, TH>::BYTES, Lds<RAWOPS_DEMOSAIC_BILINEAR, TW, TH>::BYTES);
    static constexpr int WAVES_PER_EU = LDS <= 160 * 1024 / 3 ? 6 : 2;
};

template <int TW, int TH>
__global__ void __launch_bounds__(kRenderThreads) __attribute__((amdgpu_waves_per_eu(BatchTile<TW, TH>::WAVES_PER_EU)))
batch_tiles(const BatchItem* __restrict__ items, int n, uint32_t* remaining, unsigned long long* done) {
    __shared__ __attribute__((aligned(16))) char lds[BatchTile<TW, TH>::LDS];
    const int item = (int)blockIdx.x;
    int lo = 0, hi = n - 1;                         // request owning this item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].first <= item) lo = mid; else hi = mid - 1;
    }
    const BatchItem& q = items[lo];
    RenderArgs a;
    a.decisions = nullptr;
    a.dtiles = 0;
    a.s = q.sensor;
    a.native = q.native;
    a.reduced = q.reduced;
    a.mip = q.mip;
    a.vx = q.vx; a.vy = q.vy; a.vw = q.vw; a.vh = q.vh;
    a.ox0 = q.ox0; a.oy0 = q.oy0;
    a.tiles_x = q.tiles_x;
    a.out = q.out;
    a.out8 = q.out8;
    a.ostride = q.ostride;
    a.flags = q.flags;
    a.tap = q.tap;
    int tx, ty, lx0, ly0;
    tile_origin<TW, TH>(a, item - q.first, tx, ty, lx0, ly0);
    const SensorArgs s = a.s;
    if (q.demosaic == RAWOPS_DEMOSAIC_MENON)
        render_any_tile<RAWOPS_DEMOSAIC_MENON, TW, TH, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
    else
        render_any_tile<RAWOPS_DEMOSAIC_BILINEAR, TW, TH, kRenderThreads>(a, s, lds, tx, ty, lx0, ly0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicSub(&remaining[lo], 1u) == 1u) done[lo] = wall_clock64();
    }
}

template <int TW, int TH>
void launch_tiles(const BatchItem* items, int n, int total, uint32_t* remaining, unsigned long long* done,
                  hipStream_t stream) {
    hipLaunchKernelGGL((batch_tiles<TW, TH>), dim3(total), dim3(kRenderThreads), 0, stream, items, n, remaining, done);
}

}  // namespace

// Tile shapes the autotuner chooses between.
const BatchShape kBatchShapes[kBatchShapeCount] = {{64, 32}, {64, 64}, {128, 32}};

size_t batch_state_bytes() { return sizeof(BatchState); }

unsigned long long batch_start_time(const void* state_host) {
    return static_cast<const BatchState*>(state_host)->t0;
}

hipError_t launch_batch(int shape, const BatchItem* items, int n, int total, void* state, uint32_t* remaining,
                        unsigned long long* done, hipStream_t stream) {
    if (shape < 0 || shape >= kBatchShapeCount) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1), dim3(256), 0, stream,
                       static_cast<BatchState*>(state), items, n, remaining);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || total <= 0) return e;
    switch (shape) {
    case 0: launch_tiles<64, 32>(items, n, total, remaining, done, stream); break;
    case 1: launch_tiles<64, 64>(items, n, total, remaining, done, stream); break;
    default: launch_tiles<128, 32>(items, n, total, remaining, done, stream); break;
    }
    return hipGetLastError();
}

}  // namespace rawops
